// zdx_output.cpp — the reference's two per-step output files, written from
// host copies of the simulation state:
//   write_vtk_particles       src/main.cpp:1710-1840 (legacy VTK polydata, ASCII)
//   write_particle_state_txt  src/main.cpp:1842-1870 (restartable particle blocks)
// The text is produced with the same stream formatting as the reference
// (default float formatting for VTK, fixed with 9 decimals for the state), so
// equal state gives byte-identical files.
#define ZDX_HD
#include <fstream>
#include <iomanip>

#include "zdx_geom.h"
#include "zdx_internal.h"

namespace zdx {

namespace {
inline float f32(double v) { return static_cast<float>(v); }
}  // namespace

bool write_vtk_particles(const std::string& path, int n, const HostMesh* const* mesh, const double* state,
                         const double* mass, const double* radius, const double* F, const double* T,
                         const int32_t* counts) {
    std::ofstream out(path);
    if (!out) return false;
    out << "# vtk DataFile Version 3.0\n";
    out << "zdem particles\n";
    out << "ASCII\n";
    out << "DATASET POLYDATA\n";
    size_t total = 0;
    for (int i = 0; i < n; ++i) total += mesh[i]->tris.size() / 9;
    out << "POINTS " << total * 3 << " float\n";
    for (int i = 0; i < n; ++i) {
        const double* s = state + 16 * (size_t)i;
        Qt q; q.w = s[3]; q.x = s[4]; q.y = s[5]; q.z = s[6];
        const V3 pos = v3(s[0], s[1], s[2]);
        const std::vector<double>& t = mesh[i]->tris;
        for (size_t k = 0; k < t.size(); k += 3) {
            const V3 w = add(qrotate(q, v3(t[k], t[k + 1], t[k + 2])), pos);
            out << f32(w.x) << " " << f32(w.y) << " " << f32(w.z) << "\n";
        }
    }
    out << "POLYGONS " << total << " " << total * 4 << "\n";
    for (size_t c = 0; c < total; ++c) out << "3 " << 3 * c << " " << 3 * c + 1 << " " << 3 * c + 2 << "\n";
    out << "CELL_DATA " << total << "\n";
    // one value per triangle of each particle, in particle order
    auto per_tri = [&](auto&& emit) {
        for (int i = 0; i < n; ++i) {
            const size_t nt = mesh[i]->tris.size() / 9;
            for (size_t c = 0; c < nt; ++c) emit(i);
        }
    };
    out << "SCALARS id int 1\nLOOKUP_TABLE default\n";
    per_tri([&](int i) { out << i << "\n"; });
    out << "SCALARS mass float 1\nLOOKUP_TABLE default\n";
    per_tri([&](int i) { out << f32(mass[i]) << "\n"; });
    out << "SCALARS radius float 1\nLOOKUP_TABLE default\n";
    per_tri([&](int i) { out << f32(radius[i]) << "\n"; });
    out << "SCALARS contact_count int 1\nLOOKUP_TABLE default\n";
    per_tri([&](int i) { out << counts[i] << "\n"; });
    auto vec = [&](const char* name, const double* base, int stride, int off) {
        out << "VECTORS " << name << " float\n";
        per_tri([&](int i) {
            const double* v = base + (size_t)stride * i + off;
            out << f32(v[0]) << " " << f32(v[1]) << " " << f32(v[2]) << "\n";
        });
    };
    vec("velocity", state, 16, 7);
    vec("omega", state, 16, 10);
    vec("force", F, 3, 0);
    vec("torque", T, 3, 0);
    out << "FIELD FieldData 1\n";
    out << "orientation 4 " << total << " float\n";
    per_tri([&](int i) {
        const double* s = state + 16 * (size_t)i;
        out << f32(s[3]) << " " << f32(s[4]) << " " << f32(s[5]) << " " << f32(s[6]) << "\n";
    });
    return (bool)out;
}

bool write_particle_state_txt(const std::string& path, int n, const double* state, const ParticleInit* init) {
    std::ofstream out(path);
    if (!out) return false;
    out.setf(std::ios::fixed);
    out << std::setprecision(9);
    for (int i = 0; i < n; ++i) {
        const double* s = state + 16 * (size_t)i;
        const ParticleInit& in = init[i];
        out << "particle\n";
        out << "stl = " << in.stl << "\n";
        out << "pos = " << s[0] << " " << s[1] << " " << s[2] << "\n";
        out << "vel = " << s[7] << " " << s[8] << " " << s[9] << "\n";
        out << "quat = " << s[3] << " " << s[4] << " " << s[5] << " " << s[6] << "\n";
        out << "omega = " << s[10] << " " << s[11] << " " << s[12] << "\n";
        out << "scale = " << in.scale << "\n";
        out << "density = " << in.density << "\n";
        out << "young = " << in.young << "\n";
        out << "poisson = " << in.poisson << "\n";
        out << "mu = " << in.mu << "\n";
        out << "restitution = " << in.rest << "\n";
        out << "end_particle\n\n";
    }
    return (bool)out;
}

}  // namespace zdx
