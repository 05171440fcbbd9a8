// zdx_output.cpp — the reference's two per-step output files, written from
// host copies of the simulation state:
//   write_vtk_particles       src/main.cpp:1710-1840 (legacy VTK polydata, ASCII)
//   write_particle_state_txt  src/main.cpp:1842-1870 (restartable particle blocks)
// The text is produced with the same stream formatting as the reference
// (default float formatting for VTK, fixed with 9 decimals for the state), so
// equal state gives byte-identical files.
#define ZDX_HD
#include <charconv>
#include <cmath>
#include <cstdio>
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

// ---- polyline.py's writers (polyline.py:544-649)

std::string py_float_repr(double x) {
    if (std::isnan(x)) return "nan";
    if (std::isinf(x)) return x < 0 ? "-inf" : "inf";
    if (x == 0.0) return std::signbit(x) ? "-0.0" : "0.0";
    // shortest round-trip digits d1.d2...dn e E
    char buf[64];
    auto res = std::to_chars(buf, buf + sizeof(buf), x, std::chars_format::scientific);
    std::string sci(buf, res.ptr);
    const bool neg = sci[0] == '-';
    if (neg) sci.erase(0, 1);
    const size_t epos = sci.find('e');
    std::string digits;
    for (size_t k = 0; k < epos; ++k)
        if (sci[k] != '.') digits.push_back(sci[k]);
    while (digits.size() > 1 && digits.back() == '0') digits.pop_back();
    const int e10 = std::stoi(sci.substr(epos + 1));
    const int decpt = e10 + 1;   // value = 0.d1d2...dn x 10^decpt
    const int n = (int)digits.size();
    std::string out = neg ? "-" : "";
    if (decpt <= -4 || decpt > 16) {   // exponent notation, at least two exponent digits
        out += digits[0];
        if (n > 1) { out += '.'; out += digits.substr(1); }
        const int ex = decpt - 1;
        char eb[16];
        std::snprintf(eb, sizeof(eb), "e%c%02d", ex < 0 ? '-' : '+', ex < 0 ? -ex : ex);
        out += eb;
    } else if (decpt <= 0) {
        out += "0.";
        out.append((size_t)(-decpt), '0');
        out += digits;
    } else if (decpt < n) {
        out += digits.substr(0, (size_t)decpt);
        out += '.';
        out += digits.substr((size_t)decpt);
    } else {
        out += digits;
        out.append((size_t)(decpt - n), '0');
        out += ".0";
    }
    return out;
}

namespace {
// index of p among pts (the first within 1e-9), appending it if none is.
// The distance is numpy's norm: sqrt of a running fused multiply-add sum of
// squares (the x86-64 dot kernel numpy calls), so ties at 1e-9 fall as there.
std::size_t add_point_1e9(std::vector<V3>& pts, const V3& p) {
    for (std::size_t k = 0; k < pts.size(); ++k) {
        const double dx = p.x - pts[k].x, dy = p.y - pts[k].y, dz = p.z - pts[k].z;
        if (std::sqrt(std::fma(dz, dz, std::fma(dy, dy, dx * dx))) < 1e-9) return k;
    }
    pts.push_back(p);
    return pts.size() - 1;
}
void put_point(std::ostream& f, const V3& p) {
    f << py_float_repr(p.x) << ' ' << py_float_repr(p.y) << ' ' << py_float_repr(p.z) << '\n';
}
}  // namespace

bool write_vtk_polydata_tris(const std::string& path, const double* tris, int64_t T, const std::string& comment) {
    std::vector<V3> pts;
    std::vector<std::size_t> faces;
    faces.reserve(3 * (size_t)T);
    for (int64_t t = 0; t < T; ++t)
        for (int c = 0; c < 3; ++c) {
            const double* v = tris + 9 * t + 3 * c;
            faces.push_back(add_point_1e9(pts, v3(v[0], v[1], v[2])));
        }
    std::ofstream f(path, std::ios::binary);
    if (!f) return false;
    f << "# vtk DataFile Version 3.0\n" << comment << "\nASCII\nDATASET POLYDATA\n";
    f << "POINTS " << pts.size() << " float\n";
    for (const V3& p : pts) put_point(f, p);
    f << "POLYGONS " << T << ' ' << T * 4 << '\n';
    for (int64_t t = 0; t < T; ++t)
        f << "3 " << faces[3 * t] << ' ' << faces[3 * t + 1] << ' ' << faces[3 * t + 2] << '\n';
    return (bool)f;
}

bool write_vtk_polydata_polylines(const std::string& path, const std::vector<std::vector<V3>>& loops,
                                  const std::vector<V3>& contact_points, const std::vector<char>& has_point,
                                  const std::vector<int>& loop_contact_ids) {
    std::vector<V3> pts;
    std::vector<std::vector<std::size_t>> lines;
    for (const auto& loop : loops) {
        std::vector<std::size_t> ids;
        for (const V3& p : loop) ids.push_back(add_point_1e9(pts, p));
        if (ids.size() < 2) continue;
        ids.push_back(ids[0]);   // closed for display
        lines.push_back(ids);
    }
    std::vector<std::size_t> vtx;
    for (std::size_t k = 0; k < contact_points.size(); ++k)   // entries without a point are None there
        if (k >= has_point.size() || has_point[k]) vtx.push_back(add_point_1e9(pts, contact_points[k]));
    const std::size_t n_lines = lines.size(), n_verts = vtx.size();
    std::vector<int> cell_ids;
    for (std::size_t k = 0; k < n_lines && k < loop_contact_ids.size(); ++k) cell_ids.push_back(loop_contact_ids[k]);
    if (contact_points.size() == n_lines) {
        for (std::size_t k = 0; k < n_lines && k < loop_contact_ids.size(); ++k) cell_ids.push_back(loop_contact_ids[k]);
    } else {
        for (std::size_t k = 0; k < n_verts; ++k) cell_ids.push_back(-1);
    }
    std::ofstream f(path, std::ios::binary);
    if (!f) return false;
    f << "# vtk DataFile Version 3.0\nreconstructed_contact_loops\nASCII\nDATASET POLYDATA\n";
    f << "POINTS " << pts.size() << " float\n";
    for (const V3& p : pts) put_point(f, p);
    std::size_t total = 0;
    for (const auto& ids : lines) total += 1 + ids.size();
    f << "LINES " << n_lines << ' ' << total << '\n';
    for (const auto& ids : lines) {
        f << ids.size();
        for (std::size_t id : ids) f << ' ' << id;
        f << '\n';
    }
    if (n_verts > 0) {
        f << "VERTICES " << n_verts << ' ' << n_verts * 2 << '\n';
        for (std::size_t id : vtx) f << "1 " << id << '\n';
    }
    f << "\nCELL_DATA " << n_lines + n_verts << "\nSCALARS contact_id int 1\nLOOKUP_TABLE default\n";
    for (int c : cell_ids) f << c << '\n';
    return (bool)f;
}


void polyline_transform_points(const double* in, int64_t n, const double* t, const double* q, double* out) {
    double w = q[0], x = q[1], y = q[2], z = q[3];
    const double nq = std::sqrt(std::fma(z, z, std::fma(y, y, std::fma(x, x, w * w))));
    if (nq < 1e-30) {
        w = 1.0; x = y = z = 0.0;
    } else {
        w /= nq; x /= nq; y /= nq; z /= nq;
    }
    for (int64_t i = 0; i < n; ++i) {
        const double vx = in[3 * i], vy = in[3 * i + 1], vz = in[3 * i + 2];
        const double tx = 2.0 * (y * vz - z * vy), ty = 2.0 * (z * vx - x * vz), tz = 2.0 * (x * vy - y * vx);
        out[3 * i] = vx + w * tx + (y * tz - z * ty) + t[0];
        out[3 * i + 1] = vy + w * ty + (z * tx - x * tz) + t[1];
        out[3 * i + 2] = vz + w * tz + (x * ty - y * tx) + t[2];
    }
}

namespace {
double np_dot3(const double* a, const double* b) {
    return std::fma(a[2], b[2], std::fma(a[1], b[1], a[0] * b[0]));
}
}  // namespace

bool polyline_keeps_loop(const double* pts, int64_t m, double tol) {
    if (m < 3) return false;
    double lo[3] = {pts[0], pts[1], pts[2]}, hi[3] = {pts[0], pts[1], pts[2]};
    for (int64_t i = 1; i < m; ++i)
        for (int d = 0; d < 3; ++d) {
            lo[d] = std::fmin(lo[d], pts[3 * i + d]);
            hi[d] = std::fmax(hi[d], pts[3 * i + d]);
        }
    const double e[3] = {hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]};
    return !(std::sqrt(np_dot3(e, e)) < tol * 2.0);
}

PolylineContact polyline_contact(const double* pts, int64_t m, double k) {
    PolylineContact c{};
    double Sn[3] = {0.0, 0.0, 0.0}, Gn[3] = {0.0, 0.0, 0.0};
    for (int64_t i = 0; i < m; ++i) {
        const double* p0 = pts + 3 * i;
        const double* p1 = pts + 3 * ((i + 1) % m);
        const double cr[3] = {p0[1] * p1[2] - p0[2] * p1[1], p0[2] * p1[0] - p0[0] * p1[2],
                              p0[0] * p1[1] - p0[1] * p1[0]};
        for (int d = 0; d < 3; ++d) Sn[d] += 0.5 * cr[d];
        const double dx[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
        const double s = np_dot3(p0, p1) + (1.0 / 3.0) * np_dot3(dx, dx);
        const double g = -(1.0 / 3.0) * s;
        for (int d = 0; d < 3; ++d) Gn[d] += g * dx[d];
    }
    for (int d = 0; d < 3; ++d) { c.Sn[d] = Sn[d]; c.Gn[d] = Gn[d]; }
    const double a = std::sqrt(np_dot3(Sn, Sn));
    if (a < 1e-14) {
        c.area = 0.0;
        c.has_xc0 = false;
    } else {
        c.area = a;
        c.has_xc0 = true;
        for (int d = 0; d < 3; ++d) c.nA[d] = -Sn[d] / a;
        const double* n = c.nA;
        const double cr[3] = {n[1] * Gn[2] - n[2] * Gn[1], n[2] * Gn[0] - n[0] * Gn[2], n[0] * Gn[1] - n[1] * Gn[0]};
        for (int d = 0; d < 3; ++d) c.xc0[d] = cr[d] / a;
    }
    c.Fn = 0.5 * k * c.area;
    for (int d = 0; d < 3; ++d) c.FA[d] = c.Fn * c.nA[d];
    return c;
}

}  // namespace zdx
