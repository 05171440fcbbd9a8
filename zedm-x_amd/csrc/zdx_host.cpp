// zdx_host.cpp — host-side set-up: STL loading, mesh preparation, config
// parsing and particle constants, restating src/main.cpp of
// zhl-handsome/ZEDM-X (lines cited per function) with the same FP64
// operation order (the library is built with -ffp-contract=off).
#include <cctype>
#include <cmath>
#include <cstring>
#include <fstream>
#include <map>
#include <limits>
#include <sstream>
#include <unordered_map>

#include "zdx_internal.h"

namespace zdx {

namespace {
struct Key3 {
    int64_t x, y, z;
    bool operator==(const Key3& o) const { return x == o.x && y == o.y && z == o.z; }
};
struct Key3H {
    size_t operator()(const Key3& k) const {
        uint64_t h = (uint64_t)k.x * 0x9E3779B97F4A7C15ull;
        h ^= (uint64_t)k.y * 0xC2B2AE3D27D4EB4Full + (h << 6) + (h >> 2);
        h ^= (uint64_t)k.z * 0x165667B19E3779F9ull + (h << 6) + (h >> 2);
        return (size_t)h;
    }
};
inline double d3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
inline void cross3(const double* a, const double* b, double* r) {
    r[0] = a[1] * b[2] - a[2] * b[1];
    r[1] = a[2] * b[0] - a[0] * b[2];
    r[2] = a[0] * b[1] - a[1] * b[0];
}
}  // namespace

// load_stl, src/main.cpp:308-371
bool load_stl(const std::string& path, std::vector<double>& tris) {
    tris.clear();
    std::ifstream in(path, std::ios::binary);
    if (!in) return false;
    in.seekg(0, std::ios::end);
    std::streampos size = in.tellg();
    in.seekg(0, std::ios::beg);
    if (size < 84) return false;
    char header[80];
    in.read(header, 80);
    uint32_t n = 0;
    in.read(reinterpret_cast<char*>(&n), 4);
    if (static_cast<std::streampos>(84) + static_cast<std::streampos>(n) * 50 == size) {
        tris.resize((size_t)n * 9);
        for (uint32_t i = 0; i < n; ++i) {
            float d[12];
            uint16_t attr;
            in.read(reinterpret_cast<char*>(d), 48);
            in.read(reinterpret_cast<char*>(&attr), 2);
            for (int k = 0; k < 9; ++k) tris[9 * (size_t)i + k] = d[3 + k];
        }
        return true;
    }
    in.close();
    std::ifstream ia(path);
    if (!ia) return false;
    std::string line;
    std::vector<double> vb;
    while (std::getline(ia, line)) {
        for (auto& c : line) c = static_cast<char>(std::tolower(static_cast<unsigned char>(c)));
        if (line.find("vertex") != std::string::npos) {
            std::istringstream iss(line);
            std::string tag;
            double x, y, z;
            if (iss >> tag >> x >> y >> z) { vb.push_back(x); vb.push_back(y); vb.push_back(z); }
        }
        if (line.find("endloop") != std::string::npos && vb.size() >= 9)
            tris.insert(tris.end(), vb.end() - 9, vb.end());
    }
    return !tris.empty();
}

// build_mesh, src/main.cpp:445-497, with compute_mass_properties (373-443).
// unique_vertices (287-306) keeps a vertex unless an earlier kept one lies
// within |d|^2 <= 1e-10; a spatial hash (cell 2e-5 > 1e-5 = sqrt(1e-10)) finds
// the candidates, which gives the same kept set as the reference's scan.
HostMesh build_mesh(const std::vector<double>& in, bool center) {
    HostMesh m;
    const size_t T = in.size() / 9;
    m.tris = in;
    // reference unique vertex list (only radius and bbox_diag use it)
    std::vector<double> uv;
    {
        const double cs = 2e-5;
        std::unordered_map<Key3, std::vector<int>, Key3H> grid;
        for (size_t k = 0; k < 3 * T; ++k) {
            const double* v = &in[3 * k];
            Key3 c{(int64_t)std::floor(v[0] / cs), (int64_t)std::floor(v[1] / cs), (int64_t)std::floor(v[2] / cs)};
            bool found = false;
            for (int dx = -1; dx <= 1 && !found; ++dx)
                for (int dy = -1; dy <= 1 && !found; ++dy)
                    for (int dz = -1; dz <= 1 && !found; ++dz) {
                        auto it = grid.find(Key3{c.x + dx, c.y + dy, c.z + dz});
                        if (it == grid.end()) continue;
                        for (int u : it->second) {
                            double d[3] = {v[0] - uv[3 * u], v[1] - uv[3 * u + 1], v[2] - uv[3 * u + 2]};
                            if (d3(d, d) <= 1e-10) { found = true; break; }
                        }
                    }
            if (!found) {
                grid[c].push_back((int)(uv.size() / 3));
                uv.push_back(v[0]); uv.push_back(v[1]); uv.push_back(v[2]);
            }
        }
    }
    m.n_unique_ref = (int)(uv.size() / 3);
    // compute_mass_properties
    double ds = 0.0, cs[3] = {0, 0, 0};
    double Ixx = 0, Iyy = 0, Izz = 0, Ixy = 0, Ixz = 0, Iyz = 0;
    for (size_t t = 0; t < T; ++t) {
        const double* a = &in[9 * t];
        const double* b = a + 3;
        const double* c = a + 6;
        double bc[3];
        cross3(b, c, bc);
        double det = d3(a, bc);
        ds += det;
        double sx = a[0] + b[0] + c[0], sy = a[1] + b[1] + c[1], sz = a[2] + b[2] + c[2];
        cs[0] += sx * det; cs[1] += sy * det; cs[2] += sz * det;
        double x1 = a[0], x2 = b[0], x3 = c[0], y1 = a[1], y2 = b[1], y3 = c[1], z1 = a[2], z2 = b[2], z3 = c[2];
        double f2x = x1 * x1 + x2 * x2 + x3 * x3 + x1 * x2 + x1 * x3 + x2 * x3;
        double f2y = y1 * y1 + y2 * y2 + y3 * y3 + y1 * y2 + y1 * y3 + y2 * y3;
        double f2z = z1 * z1 + z2 * z2 + z3 * z3 + z1 * z2 + z1 * z3 + z2 * z3;
        Ixx += det * (f2y + f2z);
        Iyy += det * (f2x + f2z);
        Izz += det * (f2x + f2y);
        double gxy = 2.0 * x1 * y1 + 2.0 * x2 * y2 + 2.0 * x3 * y3 + x1 * y2 + x2 * y1 + x1 * y3 + x3 * y1 + x2 * y3 + x3 * y2;
        double gxz = 2.0 * x1 * z1 + 2.0 * x2 * z2 + 2.0 * x3 * z3 + x1 * z2 + x2 * z1 + x1 * z3 + x3 * z1 + x2 * z3 + x3 * z2;
        double gyz = 2.0 * y1 * z1 + 2.0 * y2 * z2 + 2.0 * y3 * z3 + y1 * z2 + y2 * z1 + y1 * z3 + y3 * z1 + y2 * z3 + y3 * z2;
        Ixy += det * gxy;
        Ixz += det * gxz;
        Iyz += det * gyz;
    }
    double vol = 0.0, cen[3] = {0, 0, 0}, I[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (!(std::abs(ds) < 1e-18)) {
        if (ds < 0.0) {
            ds = -ds; cs[0] *= -1.0; cs[1] *= -1.0; cs[2] *= -1.0;
            Ixx *= -1.0; Iyy *= -1.0; Izz *= -1.0; Ixy *= -1.0; Ixz *= -1.0; Iyz *= -1.0;
        }
        vol = ds / 6.0;
        double den = 4.0 * ds;
        cen[0] = cs[0] / den; cen[1] = cs[1] / den; cen[2] = cs[2] / den;
        Ixx /= 60.0; Iyy /= 60.0; Izz /= 60.0;
        Ixy /= -120.0; Ixz /= -120.0; Iyz /= -120.0;
        I[0] = Ixx; I[4] = Iyy; I[8] = Izz;
        I[1] = I[3] = Ixy;
        I[2] = I[6] = Ixz;
        I[5] = I[7] = Iyz;
    }
    m.volume = vol;
    std::memcpy(m.center, cen, sizeof(cen));
    double Icm[9];
    std::memcpy(Icm, I, sizeof(I));
    if (vol > 0.0) {
        // inertia_cm = I - (|c|^2 Id - c c^T) * V  (mat3_sub/mat3_scale/mat3_outer 122-148)
        double cc = d3(cen, cen);
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                double idm = (i == j ? 1.0 : 0.0) * cc;
                double shift = idm - cen[i] * cen[j];
                Icm[3 * i + j] = I[3 * i + j] - shift * vol;
            }
    }
    std::memcpy(m.inertia, Icm, sizeof(Icm));
    if (center) {
        for (size_t k = 0; k < 3 * T; ++k) {
            m.tris[3 * k] -= cen[0]; m.tris[3 * k + 1] -= cen[1]; m.tris[3 * k + 2] -= cen[2];
        }
        for (size_t k = 0; k < uv.size() / 3; ++k) {
            uv[3 * k] -= cen[0]; uv[3 * k + 1] -= cen[1]; uv[3 * k + 2] -= cen[2];
        }
        m.center[0] = m.center[1] = m.center[2] = 0.0;
    }
    double r2 = 0.0;
    const double inf = std::numeric_limits<double>::infinity();
    double mn[3] = {inf, inf, inf}, mx[3] = {-inf, -inf, -inf};
    for (size_t k = 0; k < uv.size() / 3; ++k) {
        const double* v = &uv[3 * k];
        r2 = std::max(r2, d3(v, v));
        for (int c = 0; c < 3; ++c) { mn[c] = std::min(mn[c], v[c]); mx[c] = std::max(mx[c], v[c]); }
    }
    m.radius = std::sqrt(r2);
    double dd[3] = {mx[0] - mn[0], mx[1] - mn[1], mx[2] - mn[2]};
    m.bbox_diag = std::sqrt(d3(dd, dd));
    double es = 0.0;
    size_t ec = 0;
    for (size_t t = 0; t < T; ++t) {
        const double* v = &m.tris[9 * t];
        double e1[3] = {v[3] - v[0], v[4] - v[1], v[5] - v[2]};
        double e2[3] = {v[6] - v[3], v[7] - v[4], v[8] - v[5]};
        double e3[3] = {v[0] - v[6], v[1] - v[7], v[2] - v[8]};
        es += std::sqrt(d3(e1, e1));
        es += std::sqrt(d3(e2, e2));
        es += std::sqrt(d3(e3, e3));
        ec += 3;
    }
    if (ec > 0) m.mean_edge = es / static_cast<double>(ec);
    // Triangles whose world-frame plane test (plane_from_tri, |n| < 1e-30) is
    // not settled by the body frame.  A world corner is R(q) v + p with
    // |p| <= 1e3 (checked per body on the device); its rounding error is below
    // d = 64 eps (R + 1e3), so an edge moves by at most 2 d and |e1 x e2| by at
    // most 4 d Lmax + 4 d^2 (rotation by a unit quaternion keeps |e1 x e2| up
    // to a relative 1e-15).  Any triangle with 0.99 |e1 x e2| above that bound
    // (plus 1e-30) stays non-degenerate; the rest are tested on the device.
    {
        const double eps = 1.1102230246251565e-16;
        const double dd = 64.0 * eps * (m.radius + 1e3);
        for (size_t t = 0; t < T; ++t) {
            const double* v = &m.tris[9 * t];
            double e1[3] = {v[3] - v[0], v[4] - v[1], v[5] - v[2]};
            double e2[3] = {v[6] - v[0], v[7] - v[1], v[8] - v[2]};
            double e3[3] = {v[6] - v[3], v[7] - v[4], v[8] - v[5]};
            double c[3];
            cross3(e1, e2, c);
            double ln = std::sqrt(d3(c, c));
            double L = std::sqrt(std::max(d3(e1, e1), std::max(d3(e2, e2), d3(e3, e3))));
            if (!(0.99 * ln > 4.0 * dd * L + 4.0 * dd * dd + 1e-30)) m.suspect.push_back((int32_t)t);
        }
    }
    // indexed form: bitwise-equal corners share a slot
    {
        std::unordered_map<Key3, int, Key3H> slot;
        m.tri_idx.resize(3 * T);
        for (size_t k = 0; k < 3 * T; ++k) {
            Key3 key;
            std::memcpy(&key.x, &m.tris[3 * k], 8);
            std::memcpy(&key.y, &m.tris[3 * k + 1], 8);
            std::memcpy(&key.z, &m.tris[3 * k + 2], 8);
            auto it = slot.find(key);
            if (it == slot.end()) {
                int id = (int)(m.verts.size() / 3);
                slot.emplace(key, id);
                m.verts.push_back(m.tris[3 * k]); m.verts.push_back(m.tris[3 * k + 1]); m.verts.push_back(m.tris[3 * k + 2]);
                m.tri_idx[k] = id;
            } else {
                m.tri_idx[k] = it->second;
            }
        }
    }
    return m;
}

// parse_config_file, src/main.cpp:1531-1689
bool parse_config(const std::string& path, SimConfig& c) {
    std::ifstream in(path);
    if (!in) return false;
    std::string line;
    bool inp = false, inw = false;
    ParticleInit cp;
    WallInit cw;
    while (std::getline(in, line)) {
        if (line.empty()) continue;
        std::string cl;
        for (char ch : line) {
            if (ch == '#') break;
            cl.push_back(ch == '=' ? ' ' : ch);
        }
        std::istringstream is(cl);
        std::string k;
        if (!(is >> k)) continue;
        if (k == "particle") {
            if (inp) c.parts.push_back(cp);
            if (inw) { c.walls.push_back(cw); inw = false; }
            cp = ParticleInit{}; inp = true; continue;
        }
        if (k == "end_particle" || k == "particle_end") { if (inp) { c.parts.push_back(cp); inp = false; } continue; }
        if (k == "wall") {
            if (inp) { c.parts.push_back(cp); inp = false; }
            if (inw) c.walls.push_back(cw);
            cw = WallInit{}; inw = true; continue;
        }
        if (k == "end_wall" || k == "wall_end") { if (inw) { c.walls.push_back(cw); inw = false; } continue; }
        if (inp) {
            if (k == "stl") is >> cp.stl;
            else if (k == "pos") is >> cp.pos[0] >> cp.pos[1] >> cp.pos[2];
            else if (k == "vel") is >> cp.vel[0] >> cp.vel[1] >> cp.vel[2];
            else if (k == "omega") is >> cp.omega[0] >> cp.omega[1] >> cp.omega[2];
            else if (k == "quat") is >> cp.quat[0] >> cp.quat[1] >> cp.quat[2] >> cp.quat[3];
            else if (k == "scale") is >> cp.scale;
            else if (k == "density") is >> cp.density;
            else if (k == "young") is >> cp.young;
            else if (k == "poisson") is >> cp.poisson;
            else if (k == "mu") is >> cp.mu;
            else if (k == "restitution") is >> cp.rest;
        } else if (inw) {
            if (k == "stl") is >> cw.stl;
            else if (k == "pos") is >> cw.pos[0] >> cw.pos[1] >> cw.pos[2];
            else if (k == "quat") is >> cw.quat[0] >> cw.quat[1] >> cw.quat[2] >> cw.quat[3];
            else if (k == "scale") is >> cw.scale;
            else if (k == "mu") is >> cw.mu;
            else if (k == "restitution") is >> cw.rest;
        } else if (k == "stl") is >> c.stl;
        else if (k == "n") is >> c.n;
        else if (k == "steps") is >> c.steps;
        else if (k == "dt") is >> c.dt;
        else if (k == "spacing") is >> c.spacing;
        else if (k == "split_contacts") { int v = 1; is >> v; c.split = (v != 0); }
        else if (k == "contact_debug") { int v = 0; is >> v; c.contact_debug = (v != 0); }
        else if (k == "gravity") is >> c.gravity[0] >> c.gravity[1] >> c.gravity[2];
        else if (k == "center_mesh") { int v = 1; is >> v; c.center = (v != 0); }
        else if (k == "force_model") {
            std::string v;
            is >> v;
            c.force_model = v == "minimal" ? 1 : v == "zdem_damped" ? 2 : 0;
        }
        else if (k == "minimal_k") is >> c.k_minimal;
        else if (k == "vtk_prefix") is >> c.vtk_prefix;
        else if (k == "output_interval") is >> c.output_interval;
        else if (k == "output_dir") is >> c.output_dir;
        else if (k == "v0") is >> c.v0[0] >> c.v0[1] >> c.v0[2];
        else if (k == "v1") is >> c.v1[0] >> c.v1[1] >> c.v1[2];
    }
    if (inp) c.parts.push_back(cp);
    if (inw) c.walls.push_back(cw);
    if (c.output_interval < 1) c.output_interval = 1;
    if (!c.stl.empty()) return true;
    return !c.parts.empty();
}

bool build_scene(const SimConfig& cfg, const std::string& base_dir, ConfigScene& sc, std::string& err) {
    auto resolve = [&](const std::string& p) {
        if (base_dir.empty() || p.empty() || p[0] == '/') return p;
        std::string b = base_dir;
        if (b.back() != '/') b.push_back('/');
        return b + p;
    };
    std::map<std::string, std::vector<double>> tcache;   // per path
    std::map<std::string, int> mcache;                   // per path|scale
    auto load_for = [&](const std::string& p, double scale) -> int {
        const std::string key = p + "|" + std::to_string(scale);
        auto it = mcache.find(key);
        if (it != mcache.end()) return it->second;
        auto jt = tcache.find(p);
        if (jt == tcache.end()) {
            std::vector<double> t;
            if (!load_stl(resolve(p), t)) return -1;
            jt = tcache.emplace(p, std::move(t)).first;
        }
        std::vector<double> t = jt->second;
        for (auto& v : t) v *= scale;
        sc.meshes.push_back(build_mesh(t, cfg.center));
        const int id = (int)sc.meshes.size() - 1;
        mcache[key] = id;
        return id;
    };
    sc = ConfigScene{};
    sc.inits = cfg.parts;
    if (sc.inits.empty()) {
        sc.inits.resize(cfg.n > 0 ? cfg.n : 0);
        for (int i = 0; i < cfg.n; ++i) {
            ParticleInit& in = sc.inits[i];
            in.stl = cfg.stl;
            in.pos[0] = cfg.spacing * i;
            const double* v = (i == 0) ? cfg.v0 : (i == 1 ? cfg.v1 : nullptr);
            for (int k = 0; k < 3; ++k) in.vel[k] = v ? v[k] : 0.0;
            in.scale = 1.0; in.density = 1.0; in.young = 1e7; in.poisson = 0.25; in.mu = 0.5; in.rest = 0.5;
        }
    }
    for (const auto& in : sc.inits) {
        const std::string p = in.stl.empty() ? cfg.stl : in.stl;
        if (p.empty()) { err = "particle without stl"; return false; }
        const int id = load_for(p, in.scale > 0.0 ? in.scale : 1.0);
        if (id < 0) { err = "failed to load STL " + p; return false; }
        if (sc.meshes[id].verts.empty()) { err = "mesh has no vertices: " + p; return false; }
        sc.mesh_of.push_back(id);
    }
    // walls (main.cpp:1990-2022): a wall whose STL does not load is skipped
    for (const auto& wi : cfg.walls) {
        if (wi.stl.empty()) continue;
        std::vector<double> t;
        if (!load_stl(resolve(wi.stl), t)) continue;
        const double s = wi.scale > 0.0 ? wi.scale : 1.0;
        for (auto& v : t) v = v * s;
        sc.wall_meshes.push_back(build_mesh(t, cfg.center));
        sc.walls.push_back(wi);
    }
    return true;
}

namespace {
// mat3_det / mat3_inverse, src/main.cpp:197-220
void m_inverse(const double* a, double* r) {
    auto A = [&](int i, int j) { return a[3 * i + j]; };
    double det = A(0, 0) * (A(1, 1) * A(2, 2) - A(1, 2) * A(2, 1)) - A(0, 1) * (A(1, 0) * A(2, 2) - A(1, 2) * A(2, 0)) +
                 A(0, 2) * (A(1, 0) * A(2, 1) - A(1, 1) * A(2, 0));
    if (std::abs(det) < 1e-18) {
        for (int k = 0; k < 9; ++k) r[k] = (k % 4 == 0) ? 1.0 : 0.0;
        return;
    }
    double id = 1.0 / det;
    r[0] = (A(1, 1) * A(2, 2) - A(1, 2) * A(2, 1)) * id;
    r[1] = -(A(0, 1) * A(2, 2) - A(0, 2) * A(2, 1)) * id;
    r[2] = (A(0, 1) * A(1, 2) - A(0, 2) * A(1, 1)) * id;
    r[3] = -(A(1, 0) * A(2, 2) - A(1, 2) * A(2, 0)) * id;
    r[4] = (A(0, 0) * A(2, 2) - A(0, 2) * A(2, 0)) * id;
    r[5] = -(A(0, 0) * A(1, 2) - A(0, 2) * A(1, 0)) * id;
    r[6] = (A(1, 0) * A(2, 1) - A(1, 1) * A(2, 0)) * id;
    r[7] = -(A(0, 0) * A(2, 1) - A(0, 1) * A(2, 0)) * id;
    r[8] = (A(0, 0) * A(1, 1) - A(0, 1) * A(1, 0)) * id;
}
}  // namespace

// particle set-up, src/main.cpp:1955-1984
ParticleConst particle_const(const HostMesh& m, int mesh_id, double density, double young, double poisson, double mu,
                             double rest) {
    ParticleConst p{};
    double den = density > 0.0 ? density : 1.0;
    if (m.volume > 0.0) {
        p.mass = den * m.volume;
        p.inv_mass = 1.0 / p.mass;
        for (int k = 0; k < 9; ++k) p.Ibody[k] = m.inertia[k] * den;
    } else {
        p.mass = den;
        p.inv_mass = 1.0 / p.mass;
        double I = 0.4 * p.mass * m.radius * m.radius;
        for (int k = 0; k < 9; ++k) p.Ibody[k] = (k % 4 == 0) ? I : 0.0;
    }
    m_inverse(p.Ibody, p.Ibinv);
    p.radius = m.radius;
    p.eqr = m.volume > 0.0 ? std::cbrt(3.0 * m.volume / (4.0 * 3.141592653589793)) : m.radius;
    p.young = young;
    p.poisson = poisson;
    p.mu = mu;
    p.rest = rest;
    p.mesh = mesh_id;
    return p;
}

// L = inertia_world(p) * omega with inertia_world = R I Rt (1481-1485)
void initial_L(const ParticleConst& c, const double* q, const double* om, double* L) {
    double w0 = q[0], x0 = q[1], y0 = q[2], z0 = q[3];
    double n = std::sqrt(w0 * w0 + x0 * x0 + y0 * y0 + z0 * z0);
    double w, x, y, z;
    if (n < 1e-14) { w = 1.0; x = y = z = 0.0; }
    else { w = w0 / n; x = x0 / n; y = y0 / n; z = z0 / n; }
    double R[9];
    R[0] = 1.0 - 2.0 * (y * y + z * z); R[1] = 2.0 * (x * y - z * w); R[2] = 2.0 * (x * z + y * w);
    R[3] = 2.0 * (x * y + z * w); R[4] = 1.0 - 2.0 * (x * x + z * z); R[5] = 2.0 * (y * z - x * w);
    R[6] = 2.0 * (x * z - y * w); R[7] = 2.0 * (y * z + x * w); R[8] = 1.0 - 2.0 * (x * x + y * y);
    auto mul = [](const double* a, const double* b, double* r) {
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                double s = 0.0;
                for (int k = 0; k < 3; ++k) s += a[3 * i + k] * b[3 * k + j];
                r[3 * i + j] = s;
            }
    };
    double Rt[9], RI[9], Iw[9];
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) Rt[3 * i + j] = R[3 * j + i];
    mul(R, c.Ibody, RI);
    mul(RI, Rt, Iw);
    for (int i = 0; i < 3; ++i) L[i] = Iw[3 * i] * om[0] + Iw[3 * i + 1] * om[1] + Iw[3 * i + 2] * om[2];
}

}  // namespace zdx
