// zdx_oracle.cpp — CPU FP64 restatement of the ZDEM-X Route-B contact step.
//
// TEST INFRASTRUCTURE ONLY (the parity checker): loaded by tests/, by
// __graft_entry__.smoke() and by bench.py's cpu_baseline leg; never by the
// product.  Build: g++ -O2 -std=c++17 -ffp-contract=off (oracle/Makefile).
//
// Each function names the reference lines it restates (zhl-handsome/ZEDM-X,
// src/main.cpp unless marked polyline.py).  Operation order follows the
// reference expression by expression, so results are bit-identical where the
// reference uses only + - * / sqrt floor (all correctly rounded); cbrt, pow and
// log come from the same glibc as the reference build.
//
// Pinning (what checks each part; tests/test_oracle.py, tests/test_golden.py):
//   load_stl, build_mesh          golden vectors from the compiled reference
//                                 (harness oracle/ref_harness.cpp) + closed-form
//                                 cube/tetra volume and inertia
//   transform, tri_tri            golden vectors + hand-built crossing and
//                                 coplanar known answers
//   components, snap, T-split,    golden full-pipeline vectors (segment/comp/
//   extract_loops                 loop counters and exact loop vertex lists)
//                                 + brute-force O(n^2) components on tiny inputs
//   Sn/Gn/xc0                     golden vectors + planar-polygon closed form
//                                 Sn = A n, xc0 = -2/3 c_perp
//   sim step (forces, walls,      golden trajectories from the compiled
//   integrator, history)          reference (full-precision state dumps) +
//                                 torque-free rotation keeps L exactly
//   minimal force Fn = k/2|Sn|    closed form (polyline.py:655-701); no
//                                 reference vectors (Python is not run) ->
//                                 pinned only through the closed form.
#include "zdx_oracle.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <limits>
#include <sstream>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <vector>

namespace {

struct V3 {
    double x = 0.0, y = 0.0, z = 0.0;
    V3() = default;
    V3(double a, double b, double c) : x(a), y(b), z(c) {}
    V3 operator+(const V3& o) const { return V3{x + o.x, y + o.y, z + o.z}; }
    V3 operator-(const V3& o) const { return V3{x - o.x, y - o.y, z - o.z}; }
    V3 operator*(double s) const { return V3{x * s, y * s, z * s}; }
    V3 operator/(double s) const { return V3{x / s, y / s, z / s}; }
    V3& operator+=(const V3& o) { x += o.x; y += o.y; z += o.z; return *this; }
    V3& operator-=(const V3& o) { x -= o.x; y -= o.y; z -= o.z; return *this; }
    V3& operator*=(double s) { x *= s; y *= s; z *= s; return *this; }
};
// main.cpp:37-63
inline double dot(const V3& a, const V3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(const V3& a, const V3& b) {
    return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline double norm(const V3& v) { return std::sqrt(dot(v, v)); }
inline double norm2(const V3& v) { return dot(v, v); }
inline V3 normalize(const V3& v) {
    double n = norm(v);
    if (n < 1e-14) return V3{0.0, 0.0, 0.0};
    return v / n;
}

struct Q { double w = 1.0, x = 0.0, y = 0.0, z = 0.0; };
// main.cpp:72-98
inline Q qconj(const Q& q) { return Q{q.w, -q.x, -q.y, -q.z}; }
inline Q qmul(const Q& a, const Q& b) {
    return Q{a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z,
             a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y,
             a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x,
             a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w};
}
inline Q qnormalize(const Q& q) {
    double n = std::sqrt(q.w * q.w + q.x * q.x + q.y * q.y + q.z * q.z);
    if (n < 1e-14) return Q{};
    return Q{q.w / n, q.x / n, q.y / n, q.z / n};
}
inline V3 qrotate(const Q& q, const V3& v) {
    Q p{0.0, v.x, v.y, v.z};
    Q qn = qnormalize(q);
    Q r = qmul(qmul(qn, p), qconj(qn));
    return V3{r.x, r.y, r.z};
}

struct M3 { double m[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}; };
// main.cpp:104-220
inline M3 m_identity() { M3 r; r.m[0][0] = r.m[1][1] = r.m[2][2] = 1.0; return r; }
inline M3 m_sub(const M3& a, const M3& b) {
    M3 r; for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) r.m[i][j] = a.m[i][j] - b.m[i][j]; return r;
}
inline M3 m_scale(const M3& a, double s) {
    M3 r; for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) r.m[i][j] = a.m[i][j] * s; return r;
}
inline M3 m_outer(const V3& a, const V3& b) {
    M3 r;
    r.m[0][0] = a.x * b.x; r.m[0][1] = a.x * b.y; r.m[0][2] = a.x * b.z;
    r.m[1][0] = a.y * b.x; r.m[1][1] = a.y * b.y; r.m[1][2] = a.y * b.z;
    r.m[2][0] = a.z * b.x; r.m[2][1] = a.z * b.y; r.m[2][2] = a.z * b.z;
    return r;
}
inline M3 q_to_m(const Q& q) {
    Q n = qnormalize(q);
    double w = n.w, x = n.x, y = n.y, z = n.z;
    M3 R;
    R.m[0][0] = 1.0 - 2.0 * (y * y + z * z);
    R.m[0][1] = 2.0 * (x * y - z * w);
    R.m[0][2] = 2.0 * (x * z + y * w);
    R.m[1][0] = 2.0 * (x * y + z * w);
    R.m[1][1] = 1.0 - 2.0 * (x * x + z * z);
    R.m[1][2] = 2.0 * (y * z - x * w);
    R.m[2][0] = 2.0 * (x * z - y * w);
    R.m[2][1] = 2.0 * (y * z + x * w);
    R.m[2][2] = 1.0 - 2.0 * (x * x + y * y);
    return R;
}
inline V3 m_mulv(const M3& m, const V3& v) {
    return V3{m.m[0][0] * v.x + m.m[0][1] * v.y + m.m[0][2] * v.z,
              m.m[1][0] * v.x + m.m[1][1] * v.y + m.m[1][2] * v.z,
              m.m[2][0] * v.x + m.m[2][1] * v.y + m.m[2][2] * v.z};
}
inline M3 m_mul(const M3& a, const M3& b) {
    M3 r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            r.m[i][j] = 0.0;
            for (int k = 0; k < 3; ++k) r.m[i][j] += a.m[i][k] * b.m[k][j];
        }
    return r;
}
inline M3 m_T(const M3& a) { M3 r; for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) r.m[i][j] = a.m[j][i]; return r; }
inline double m_det(const M3& a) {
    return a.m[0][0] * (a.m[1][1] * a.m[2][2] - a.m[1][2] * a.m[2][1])
         - a.m[0][1] * (a.m[1][0] * a.m[2][2] - a.m[1][2] * a.m[2][0])
         + a.m[0][2] * (a.m[1][0] * a.m[2][1] - a.m[1][1] * a.m[2][0]);
}
inline M3 m_inv(const M3& a) {
    double det = m_det(a);
    if (std::abs(det) < 1e-18) return m_identity();
    double id = 1.0 / det;
    M3 r;
    r.m[0][0] =  (a.m[1][1] * a.m[2][2] - a.m[1][2] * a.m[2][1]) * id;
    r.m[0][1] = -(a.m[0][1] * a.m[2][2] - a.m[0][2] * a.m[2][1]) * id;
    r.m[0][2] =  (a.m[0][1] * a.m[1][2] - a.m[0][2] * a.m[1][1]) * id;
    r.m[1][0] = -(a.m[1][0] * a.m[2][2] - a.m[1][2] * a.m[2][0]) * id;
    r.m[1][1] =  (a.m[0][0] * a.m[2][2] - a.m[0][2] * a.m[2][0]) * id;
    r.m[1][2] = -(a.m[0][0] * a.m[1][2] - a.m[0][2] * a.m[1][0]) * id;
    r.m[2][0] =  (a.m[1][0] * a.m[2][1] - a.m[1][1] * a.m[2][0]) * id;
    r.m[2][1] = -(a.m[0][0] * a.m[2][1] - a.m[0][1] * a.m[2][0]) * id;
    r.m[2][2] =  (a.m[0][0] * a.m[1][1] - a.m[0][1] * a.m[1][0]) * id;
    return r;
}

using Tri = std::array<V3, 3>;

struct Mesh {
    std::vector<V3> verts;
    std::vector<Tri> tris;
    double radius = 0, mean_edge = 0, bbox_diag = 0, volume = 0;
    V3 center;
    M3 inertia_unit;
};

// main.cpp:287-306
std::vector<V3> unique_vertices(const std::vector<Tri>& tris) {
    std::vector<V3> out;
    for (const auto& t : tris)
        for (const auto& v : t) {
            bool found = false;
            for (const auto& u : out) {
                V3 d = v - u;
                if (dot(d, d) <= 1e-10) { found = true; break; }
            }
            if (!found) out.push_back(v);
        }
    return out;
}

// main.cpp:308-371
bool load_stl(const std::string& path, std::vector<Tri>& tris) {
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
        tris.reserve(n);
        for (uint32_t i = 0; i < n; ++i) {
            float d[12];
            uint16_t attr;
            in.read(reinterpret_cast<char*>(d), 48);
            in.read(reinterpret_cast<char*>(&attr), 2);
            tris.push_back(Tri{V3{d[3], d[4], d[5]}, V3{d[6], d[7], d[8]}, V3{d[9], d[10], d[11]}});
        }
        return true;
    }
    in.close();
    std::ifstream ia(path);
    if (!ia) return false;
    std::string line;
    std::vector<V3> vb;
    while (std::getline(ia, line)) {
        for (auto& c : line) c = static_cast<char>(std::tolower(static_cast<unsigned char>(c)));
        if (line.find("vertex") != std::string::npos) {
            std::istringstream iss(line);
            std::string tag;
            double x, y, z;
            if (iss >> tag >> x >> y >> z) vb.emplace_back(x, y, z);
        }
        if (line.find("endloop") != std::string::npos && vb.size() >= 3)
            tris.push_back(Tri{vb[vb.size() - 3], vb[vb.size() - 2], vb[vb.size() - 1]});
    }
    return !tris.empty();
}

// main.cpp:373-443
void mass_props(const std::vector<Tri>& tris, double& vol, V3& cen, M3& I) {
    double ds = 0.0;
    V3 cs;
    double Ixx = 0, Iyy = 0, Izz = 0, Ixy = 0, Ixz = 0, Iyz = 0;
    for (const auto& t : tris) {
        V3 a = t[0], b = t[1], c = t[2];
        double det = dot(a, cross(b, c));
        ds += det;
        cs += (a + b + c) * det;
        double x1 = a.x, x2 = b.x, x3 = c.x, y1 = a.y, y2 = b.y, y3 = c.y, z1 = a.z, z2 = b.z, z3 = c.z;
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
    if (std::abs(ds) < 1e-18) { vol = 0.0; cen = V3{}; I = M3{}; return; }
    if (ds < 0.0) {
        ds = -ds; cs *= -1.0;
        Ixx *= -1.0; Iyy *= -1.0; Izz *= -1.0; Ixy *= -1.0; Ixz *= -1.0; Iyz *= -1.0;
    }
    vol = ds / 6.0;
    cen = cs / (4.0 * ds);
    Ixx /= 60.0; Iyy /= 60.0; Izz /= 60.0;
    Ixy /= -120.0; Ixz /= -120.0; Iyz /= -120.0;
    M3 r;
    r.m[0][0] = Ixx; r.m[1][1] = Iyy; r.m[2][2] = Izz;
    r.m[0][1] = r.m[1][0] = Ixy;
    r.m[0][2] = r.m[2][0] = Ixz;
    r.m[1][2] = r.m[2][1] = Iyz;
    I = r;
}

// main.cpp:445-497
Mesh build_mesh(const std::vector<Tri>& tris, bool center) {
    Mesh m;
    m.verts = unique_vertices(tris);
    m.tris = tris;
    V3 c;
    M3 Io;
    mass_props(m.tris, m.volume, c, Io);
    m.center = c;
    M3 Icm = Io;
    if (m.volume > 0.0) {
        M3 shift = m_sub(m_scale(m_identity(), dot(c, c)), m_outer(c, c));
        Icm = m_sub(Io, m_scale(shift, m.volume));
    }
    if (center) {
        for (auto& v : m.verts) v -= c;
        for (auto& t : m.tris) { t[0] -= c; t[1] -= c; t[2] -= c; }
        m.center = V3{};
    }
    m.inertia_unit = Icm;
    double r2 = 0.0;
    const double inf = std::numeric_limits<double>::infinity();
    V3 mn{inf, inf, inf}, mx{-inf, -inf, -inf};
    for (const auto& v : m.verts) {
        r2 = std::max(r2, dot(v, v));
        mn.x = std::min(mn.x, v.x); mn.y = std::min(mn.y, v.y); mn.z = std::min(mn.z, v.z);
        mx.x = std::max(mx.x, v.x); mx.y = std::max(mx.y, v.y); mx.z = std::max(mx.z, v.z);
    }
    m.radius = std::sqrt(r2);
    m.bbox_diag = norm(mx - mn);
    double es = 0.0;
    std::size_t ec = 0;
    for (const auto& t : m.tris) {
        es += norm(t[1] - t[0]);
        es += norm(t[2] - t[1]);
        es += norm(t[0] - t[2]);
        ec += 3;
    }
    if (ec > 0) m.mean_edge = es / static_cast<double>(ec);
    return m;
}

struct Pose { V3 pos; Q rot; };

// main.cpp:499-510
std::vector<Tri> transform_tris(const Mesh& m, const Pose& p) {
    std::vector<Tri> out;
    out.reserve(m.tris.size());
    for (const auto& t : m.tris)
        out.push_back(Tri{qrotate(p.rot, t[0]) + p.pos, qrotate(p.rot, t[1]) + p.pos, qrotate(p.rot, t[2]) + p.pos});
    return out;
}
// main.cpp:512-516
V3 tri_normal(const Tri& t) { return normalize(cross(t[1] - t[0], t[2] - t[0])); }
// main.cpp:518-527
bool plane_from_tri(const Tri& t, V3& n, double& d) {
    n = cross(t[1] - t[0], t[2] - t[0]);
    double ln = norm(n);
    if (ln < 1e-30) return false;
    n = n / ln;
    d = -dot(n, t[0]);
    return true;
}
// main.cpp:529-543
bool point_in_tri(const V3& p, const Tri& t, const V3& n) {
    V3 a = t[0], b = t[1], c = t[2];
    V3 ab = b - a, bc = c - b, ca = a - c;
    V3 ap = p - a, bp = p - b, cp = p - c;
    double c1 = dot(cross(ab, ap), n);
    double c2 = dot(cross(bc, bp), n);
    double c3 = dot(cross(ca, cp), n);
    return (c1 >= -1e-10) && (c2 >= -1e-10) && (c3 >= -1e-10);
}
// main.cpp:545-559
bool seg_plane(const V3& p0, const V3& p1, const V3& n, double d, V3& out) {
    V3 u = p1 - p0;
    double den = dot(n, u);
    double num = -(dot(n, p0) + d);
    if (std::abs(den) < 1e-30) return false;
    double t = num / den;
    if (t < -1e-10 || t > 1.0 + 1e-10) return false;
    t = std::min(1.0, std::max(0.0, t));
    out = p0 + u * t;
    return true;
}
// main.cpp:561-577
std::vector<V3> unique_points(const std::vector<V3>& pts, double tol) {
    std::vector<V3> out;
    double t2 = tol * tol;
    for (const auto& p : pts) {
        bool keep = true;
        for (const auto& q : out)
            if (norm2(p - q) <= t2) { keep = false; break; }
        if (keep) out.push_back(p);
    }
    return out;
}
// main.cpp:579-631
bool tri_tri(const Tri& A, const V3& nA, double dA, const Tri& B, const V3& nB, double dB, V3& o0, V3& o1) {
    V3 tau = cross(nA, nB);
    double lt = norm(tau);
    if (lt < 1e-20) return false;
    tau = tau / lt;
    std::vector<V3> c;
    for (int i = 0; i < 3; ++i) {
        V3 p;
        if (seg_plane(A[i], A[(i + 1) % 3], nB, dB, p) && point_in_tri(p, B, nB)) c.push_back(p);
    }
    for (int i = 0; i < 3; ++i) {
        V3 p;
        if (seg_plane(B[i], B[(i + 1) % 3], nA, dA, p) && point_in_tri(p, A, nA)) c.push_back(p);
    }
    c = unique_points(c, 1e-9);
    if (c.size() < 2) return false;
    double mn = std::numeric_limits<double>::infinity(), mx = -std::numeric_limits<double>::infinity();
    int i0 = -1, i1 = -1;
    for (int i = 0; i < static_cast<int>(c.size()); ++i) {
        double t = dot(c[i], tau);
        if (t < mn) { mn = t; i0 = i; }
        if (t > mx) { mx = t; i1 = i; }
    }
    if (i0 == i1) return false;
    V3 p0 = c[i0], p1 = c[i1];
    if (norm2(p1 - p0) < 1e-18) return false;
    if (dot(p1 - p0, tau) < 0.0) std::swap(p0, p1);
    o0 = p0; o1 = p1;
    return true;
}

struct I3 {
    int x = 0, y = 0, z = 0;
    bool operator==(const I3& o) const { return x == o.x && y == o.y && z == o.z; }
};
struct I3H {
    std::size_t operator()(const I3& k) const {
        return std::hash<int>{}(k.x) ^ (std::hash<int>{}(k.y) << 1) ^ (std::hash<int>{}(k.z) << 2);
    }
};
// main.cpp:649-655
inline I3 cell_of(const V3& p, double h) {
    return I3{static_cast<int>(std::floor(p.x / h)), static_cast<int>(std::floor(p.y / h)),
              static_cast<int>(std::floor(p.z / h))};
}
// main.cpp:657-678
struct UF {
    std::vector<int> par, rk;
    explicit UF(int n) : par(n), rk(n, 0) { for (int i = 0; i < n; ++i) par[i] = i; }
    int find(int x) { if (par[x] != x) par[x] = find(par[x]); return par[x]; }
    void unite(int a, int b) {
        int ra = find(a), rb = find(b);
        if (ra == rb) return;
        if (rk[ra] < rk[rb]) par[ra] = rb;
        else if (rk[ra] > rk[rb]) par[rb] = ra;
        else { par[rb] = ra; rk[ra] += 1; }
    }
};
using Seg = std::pair<V3, V3>;
using Edge = std::pair<int, int>;

// main.cpp:680-746
void snap(const std::vector<Seg>& segs, double tol, std::vector<V3>& V, std::vector<Edge>& E) {
    double h = tol;
    int ns = static_cast<int>(segs.size());
    std::vector<V3> pts(2 * ns);
    for (int i = 0; i < ns; ++i) { pts[2 * i] = segs[i].first; pts[2 * i + 1] = segs[i].second; }
    UF uf(2 * ns);
    std::unordered_map<I3, std::vector<int>, I3H> grid;
    for (int i = 0; i < 2 * ns; ++i) {
        I3 c = cell_of(pts[i], h);
        for (int dx = -1; dx <= 1; ++dx)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dz = -1; dz <= 1; ++dz) {
                    auto it = grid.find(I3{c.x + dx, c.y + dy, c.z + dz});
                    if (it == grid.end()) continue;
                    for (int j : it->second)
                        if (norm2(pts[i] - pts[j]) <= h * h) uf.unite(i, j);
                }
        grid[c].push_back(i);
    }
    std::unordered_map<int, int> vid_of;
    std::unordered_map<int, V3> sums;
    std::unordered_map<int, int> cnt;
    for (int i = 0; i < 2 * ns; ++i) {
        int r = uf.find(i);
        if (vid_of.find(r) == vid_of.end()) {
            int v = static_cast<int>(vid_of.size());
            vid_of[r] = v; sums[r] = pts[i]; cnt[r] = 1;
        } else { sums[r] += pts[i]; cnt[r] += 1; }
    }
    V.assign(vid_of.size(), V3{});
    for (const auto& kv : vid_of) V[kv.second] = sums[kv.first] / static_cast<double>(cnt[kv.first]);
    E.clear();
    for (int s = 0; s < ns; ++s) {
        int a = vid_of[uf.find(2 * s)], b = vid_of[uf.find(2 * s + 1)];
        if (a != b) E.emplace_back(a, b);
    }
}
// main.cpp:748-759
std::pair<double, double> pt_seg(const V3& p, const V3& a, const V3& b) {
    V3 ab = b - a;
    double den = dot(ab, ab);
    if (den <= 1e-12) return {norm2(p - a), 0.0};
    double t = dot(p - a, ab) / den;
    double tc = std::min(1.0, std::max(0.0, t));
    V3 pr = a + ab * tc;
    return {norm2(p - pr), tc};
}
// main.cpp:761-802
double seg_seg(const V3& p1, const V3& q1, const V3& p2, const V3& q2) {
    V3 d1 = q1 - p1, d2 = q2 - p2, r = p1 - p2;
    double a = dot(d1, d1), e = dot(d2, d2), f = dot(d2, r);
    if (a <= 1e-12 && e <= 1e-12) return dot(p1 - p2, p1 - p2);
    if (a <= 1e-12) {
        double t = std::min(1.0, std::max(0.0, f / e));
        V3 c = p2 + d2 * t;
        return dot(p1 - c, p1 - c);
    }
    if (e <= 1e-12) {
        double s = std::min(1.0, std::max(0.0, -dot(d1, r) / a));
        V3 c = p1 + d1 * s;
        return dot(c - p2, c - p2);
    }
    double b = dot(d1, d2), c = dot(d1, r);
    double den = a * e - b * b;
    double s = 0.0;
    if (std::abs(den) > 0.0) s = std::min(1.0, std::max(0.0, (b * f - c * e) / den));
    double t = (b * s + f) / e;
    if (t < 0.0) { t = 0.0; s = std::min(1.0, std::max(0.0, -c / a)); }
    else if (t > 1.0) { t = 1.0; s = std::min(1.0, std::max(0.0, (b - c) / a)); }
    V3 df = (p1 + d1 * s) - (p2 + d2 * t);
    return dot(df, df);
}
// main.cpp:804-898
void components(const std::vector<Seg>& segs, double tol, std::vector<std::vector<int>>& comps) {
    double h = tol;
    int ns = static_cast<int>(segs.size());
    comps.clear();
    if (ns == 0) return;
    std::unordered_map<I3, std::vector<int>, I3H> grid;
    std::vector<std::pair<V3, V3>> box;
    std::vector<std::array<V3, 3>> rep;
    for (int i = 0; i < ns; ++i) {
        V3 p0 = segs[i].first, p1 = segs[i].second;
        V3 mid = (p0 + p1) * 0.5;
        rep.push_back({p0, p1, mid});
        box.push_back({V3{std::min(p0.x, p1.x) - h, std::min(p0.y, p1.y) - h, std::min(p0.z, p1.z) - h},
                       V3{std::max(p0.x, p1.x) + h, std::max(p0.y, p1.y) + h, std::max(p0.z, p1.z) + h}});
        for (const auto& rp : rep.back()) grid[cell_of(rp, h)].push_back(i);
    }
    auto near = [&](int i) {
        std::vector<int> cand;
        std::vector<char> mark(ns, 0);
        for (const auto& rp : rep[i]) {
            I3 c = cell_of(rp, h);
            for (int dx = -1; dx <= 1; ++dx)
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dz = -1; dz <= 1; ++dz) {
                        auto it = grid.find(I3{c.x + dx, c.y + dy, c.z + dz});
                        if (it == grid.end()) continue;
                        for (int j : it->second) {
                            if (j == i) continue;
                            if (!mark[j]) { mark[j] = 1; cand.push_back(j); }
                        }
                    }
        }
        return cand;
    };
    auto ovl = [&](int i, int j) {
        const V3 &a0 = box[i].first, &a1 = box[i].second, &b0 = box[j].first, &b1 = box[j].second;
        return !(a1.x < b0.x || a0.x > b1.x || a1.y < b0.y || a0.y > b1.y || a1.z < b0.z || a0.z > b1.z);
    };
    double t2 = h * h;
    std::vector<char> vis(ns, 0);
    for (int seed = 0; seed < ns; ++seed) {
        if (vis[seed]) continue;
        std::vector<int> st{seed};
        vis[seed] = 1;
        std::vector<int> comp{seed};
        while (!st.empty()) {
            int i = st.back();
            st.pop_back();
            for (int j : near(i)) {
                if (vis[j]) continue;
                if (!ovl(i, j)) continue;
                if (seg_seg(segs[i].first, segs[i].second, segs[j].first, segs[j].second) <= t2) {
                    vis[j] = 1; st.push_back(j); comp.push_back(j);
                }
            }
        }
        comps.push_back(comp);
    }
    std::sort(comps.begin(), comps.end(), [](const auto& a, const auto& b) { return a.size() > b.size(); });
}
// main.cpp:900-961
void tsplit(const std::vector<V3>& V, const std::vector<Edge>& E, double tol, std::vector<Edge>& out) {
    double h = tol, t2 = h * h;
    std::unordered_map<I3, std::vector<int>, I3H> g;
    for (int v = 0; v < static_cast<int>(V.size()); ++v) g[cell_of(V[v], h)].push_back(v);
    out.clear();
    for (const auto& e : E) {
        int a = e.first, b = e.second;
        V3 pa = V[a], pb = V[b];
        V3 mn{std::min(pa.x, pb.x) - h, std::min(pa.y, pb.y) - h, std::min(pa.z, pb.z) - h};
        V3 mx{std::max(pa.x, pb.x) + h, std::max(pa.y, pb.y) + h, std::max(pa.z, pb.z) + h};
        I3 c0 = cell_of(mn, h), c1 = cell_of(mx, h);
        std::vector<std::pair<double, int>> sp;
        for (int ix = c0.x; ix <= c1.x; ++ix)
            for (int iy = c0.y; iy <= c1.y; ++iy)
                for (int iz = c0.z; iz <= c1.z; ++iz) {
                    auto it = g.find(I3{ix, iy, iz});
                    if (it == g.end()) continue;
                    for (int v : it->second) {
                        if (v == a || v == b) continue;
                        auto dt = pt_seg(V[v], pa, pb);
                        if (dt.first <= t2 && dt.second > 1e-6 && dt.second < 1.0 - 1e-6) sp.emplace_back(dt.second, v);
                    }
                }
        if (sp.empty()) { out.push_back(e); continue; }
        std::sort(sp.begin(), sp.end(), [](const auto& x, const auto& y) { return x.first < y.first; });
        std::vector<int> ch{a};
        int last = -1;
        for (const auto& s : sp) { if (s.second == last) continue; ch.push_back(s.second); last = s.second; }
        ch.push_back(b);
        for (int i = 0; i + 1 < static_cast<int>(ch.size()); ++i)
            if (ch[i] != ch[i + 1]) out.emplace_back(ch[i], ch[i + 1]);
    }
}
// main.cpp:977-982 comparator
void start_order(const std::vector<int>& deg, std::vector<int>& order) {
    int n = static_cast<int>(deg.size());
    order.resize(n);
    for (int i = 0; i < n; ++i) order[i] = i;
    std::sort(order.begin(), order.end(), [&](int i, int j) {
        int di = deg[i], dj = deg[j];
        if ((di == 2) != (dj == 2)) return di == 2;
        return di > dj;
    });
}
// main.cpp:963-1025
std::vector<std::vector<int>> loops_of(int nv, const std::vector<Edge>& E) {
    std::vector<std::vector<std::pair<int, int>>> adj(nv);
    for (int ei = 0; ei < static_cast<int>(E.size()); ++ei) {
        adj[E[ei].first].push_back({E[ei].second, ei});
        adj[E[ei].second].push_back({E[ei].first, ei});
    }
    std::vector<char> used(E.size(), 0);
    std::vector<int> deg(nv), order;
    for (int i = 0; i < nv; ++i) deg[i] = static_cast<int>(adj[i].size());
    start_order(deg, order);
    std::vector<std::vector<int>> loops;
    for (int sv : order) {
        for (const auto& nb : adj[sv]) {
            if (used[nb.second]) continue;
            std::vector<int> path{sv};
            int prev = sv, cur = nb.first;
            used[nb.second] = 1;
            path.push_back(cur);
            int steps = 0;
            while (steps < 100000) {
                steps += 1;
                if (cur == sv) {
                    if (path.size() > 1) { path.pop_back(); loops.push_back(path); }
                    break;
                }
                std::vector<std::pair<int, int>> cand;
                for (const auto& e : adj[cur]) if (!used[e.second]) cand.push_back(e);
                if (cand.empty()) break;
                auto pick = cand[0];
                for (const auto& c : cand) if (c.first != prev) { pick = c; break; }
                used[pick.second] = 1;
                prev = cur;
                cur = pick.first;
                path.push_back(cur);
            }
        }
    }
    return loops;
}
// main.cpp:1027-1040
void sn_gn(const std::vector<V3>& L, V3& Sn, V3& Gn) {
    Sn = V3{}; Gn = V3{};
    int m = static_cast<int>(L.size());
    for (int i = 0; i < m; ++i) {
        const V3& p0 = L[i];
        const V3& p1 = L[(i + 1) % m];
        Sn += cross(p0, p1) * 0.5;
        V3 dx = p1 - p0;
        double s = dot(p0, p1) + (1.0 / 3.0) * dot(dx, dx);
        Gn += dx * (-(1.0 / 3.0) * s);
    }
}
// main.cpp:1042-1052
bool xc0_of(const V3& Sn, const V3& Gn, V3& xc, V3& nA, double& area) {
    area = norm(Sn);
    if (area < 1e-14) { xc = V3{}; nA = V3{}; return false; }
    nA = Sn * (-1.0 / area);
    xc = cross(nA, Gn) / area;
    return true;
}

// ---- wall path: main.cpp:1057-1258 ----
struct V2 { double x = 0, y = 0; };
inline V2 v2sub(const V2& a, const V2& b) { return V2{a.x - b.x, a.y - b.y}; }
inline double cr2(const V2& a, const V2& b) { return a.x * b.y - a.y * b.x; }
inline bool inside(const V2& p, const V2& c1, const V2& c2) { return cr2(v2sub(c2, c1), v2sub(p, c1)) >= 0.0; }
inline V2 isect(const V2& s, const V2& e, const V2& c1, const V2& c2) {
    V2 dc = v2sub(c1, c2), dp = v2sub(s, e);
    double n1 = cr2(c1, c2), n2 = cr2(s, e), n3 = cr2(dc, dp);
    if (std::abs(n3) < 1e-30) return s;
    return V2{(n1 * dp.x - n2 * dc.x) / n3, (n1 * dp.y - n2 * dc.y) / n3};
}
std::vector<V2> sh_clip(const std::vector<V2>& subj, const std::vector<V2>& clip) {
    std::vector<V2> out = subj;
    if (clip.size() < 3 || subj.size() < 3) return out;
    V2 c1 = clip.back();
    for (const auto& c2 : clip) {
        if (out.empty()) break;
        std::vector<V2> in = out;
        out.clear();
        V2 s = in.back();
        for (const auto& e : in) {
            if (inside(e, c1, c2)) {
                if (!inside(s, c1, c2)) out.push_back(isect(s, e, c1, c2));
                out.push_back(e);
            } else if (inside(s, c1, c2)) {
                out.push_back(isect(s, e, c1, c2));
            }
            s = e;
        }
        c1 = c2;
    }
    return out;
}
double poly_area(const std::vector<V2>& p) {
    if (p.size() < 3) return 0.0;
    double a = 0.0;
    int n = static_cast<int>(p.size());
    for (int i = 0; i < n; ++i) a += cr2(p[i], p[(i + 1) % n]);
    return std::abs(a) * 0.5;
}
V2 poly_centroid(const std::vector<V2>& p) {
    if (p.empty()) return V2{};
    V2 c;
    for (const auto& q : p) { c.x += q.x; c.y += q.y; }
    double n = static_cast<double>(p.size());
    return V2{c.x / n, c.y / n};
}
void basis(const V3& n, V3& u, V3& v) {
    if (std::abs(n.x) < 0.9) u = normalize(cross(n, V3{1.0, 0.0, 0.0}));
    else u = normalize(cross(n, V3{0.0, 1.0, 0.0}));
    v = cross(n, u);
}
std::vector<V3> section(const Mesh& m, const Pose& p, const V3& pn, double pd, double tol) {
    std::vector<V3> pts;
    auto tris = transform_tris(m, p);
    for (const auto& t : tris)
        for (int i = 0; i < 3; ++i) {
            V3 p0 = t[i], p1 = t[(i + 1) % 3];
            double d0 = dot(pn, p0) + pd, d1 = dot(pn, p1) + pd;
            if (d0 * d1 < 0.0) {
                double tt = d0 / (d0 - d1);
                tt = std::min(1.0, std::max(0.0, tt));
                pts.push_back(p0 + (p1 - p0) * tt);
            }
        }
    return unique_points(pts, tol);
}
bool wall_contact(const Mesh& m, const Pose& p, const Tri& wt, const V3& wn, double tol,
                  double& area, V3& xc, V3& nrm) {
    V3 pn = wn;
    double pd = -dot(pn, wt[0]);
    auto s3 = section(m, p, pn, pd, tol);
    if (s3.size() < 3) return false;
    V3 u, v;
    basis(pn, u, v);
    V3 o = wt[0];
    std::vector<V2> subj, clip;
    for (const auto& q : s3) { V3 d = q - o; subj.push_back(V2{dot(d, u), dot(d, v)}); }
    for (const auto& q : wt) { V3 d = q - o; clip.push_back(V2{dot(d, u), dot(d, v)}); }
    auto cl = sh_clip(subj, clip);
    if (cl.size() < 3) return false;
    double a = poly_area(cl);
    if (a < tol * tol * 0.01) return false;
    V2 c = poly_centroid(cl);
    area = a;
    xc = o + u * c.x + v * c.y;
    nrm = pn;
    return true;
}

// ---- per-pair pipeline: main.cpp:2061-2219 ----
struct LoopOut {
    V3 Sn, Gn, nA, xc;
    double area;
    int comp;
    int raw_index;
    std::vector<V3> pts;
};
struct PairOut {
    int64_t tri_checks = 0;
    std::vector<Seg> segs;
    int comps = 0;
    int loops_raw = 0;
    std::vector<LoopOut> loops;   // accepted, after flip toward B
};

double tol_of(const Mesh& A) {
    double tol = A.mean_edge > 0.0 ? (A.mean_edge * 0.1) : (A.bbox_diag * 1e-2);
    if (tol <= 0.0) tol = 1e-6;
    return tol;
}

void pair_pipeline(const Mesh& mA, const Pose& pA, const Mesh& mB, const Pose& pB, bool split, PairOut& out) {
    auto tA = transform_tris(mA, pA);
    auto tB = transform_tris(mB, pB);
    int TA = static_cast<int>(tA.size()), TB = static_cast<int>(tB.size());
    std::vector<V3> nA(TA), nB(TB);
    std::vector<double> dA(TA), dB(TB);
    for (int i = 0; i < TA; ++i) { V3 n; double d; if (plane_from_tri(tA[i], n, d)) { nA[i] = n; dA[i] = d; } else { nA[i] = V3{}; dA[i] = 0.0; } }
    for (int i = 0; i < TB; ++i) { V3 n; double d; if (plane_from_tri(tB[i], n, d)) { nB[i] = n; dB[i] = d; } else { nB[i] = V3{}; dB[i] = 0.0; } }
    auto bx = [](const Tri& t, V3& mn, V3& mx) {
        mn = V3{std::min({t[0].x, t[1].x, t[2].x}), std::min({t[0].y, t[1].y, t[2].y}), std::min({t[0].z, t[1].z, t[2].z})};
        mx = V3{std::max({t[0].x, t[1].x, t[2].x}), std::max({t[0].y, t[1].y, t[2].y}), std::max({t[0].z, t[1].z, t[2].z})};
    };
    std::vector<V3> a0(TA), a1(TA), b0(TB), b1(TB);
    for (int i = 0; i < TA; ++i) bx(tA[i], a0[i], a1[i]);
    for (int i = 0; i < TB; ++i) bx(tB[i], b0[i], b1[i]);
    double tol = tol_of(mA);
    out.segs.clear();
    out.tri_checks = 0;
    for (int i = 0; i < TA; ++i) {
        if (norm2(nA[i]) < 1e-30) continue;
        for (int j = 0; j < TB; ++j) {
            out.tri_checks++;
            if (a1[i].x < b0[j].x || a0[i].x > b1[j].x || a1[i].y < b0[j].y || a0[i].y > b1[j].y ||
                a1[i].z < b0[j].z || a0[i].z > b1[j].z) continue;
            if (norm2(nB[j]) < 1e-30) continue;
            V3 p0, p1;
            if (tri_tri(tA[i], nA[i], dA[i], tB[j], nB[j], dB[j], p0, p1)) out.segs.emplace_back(p0, p1);
        }
    }
    out.comps = 0;
    out.loops_raw = 0;
    out.loops.clear();
    if (out.segs.empty()) return;
    std::vector<std::vector<int>> comps;
    if (split) components(out.segs, tol, comps);
    else {
        comps.resize(1);
        for (int s = 0; s < static_cast<int>(out.segs.size()); ++s) comps[0].push_back(s);
    }
    out.comps = static_cast<int>(comps.size());
    double area_eps = mA.mean_edge > 0.0 ? (mA.mean_edge * mA.mean_edge * 1e-4) : 1e-12;
    for (int ci = 0; ci < static_cast<int>(comps.size()); ++ci) {
        std::vector<Seg> cs;
        for (int idx : comps[ci]) cs.push_back(out.segs[idx]);
        std::vector<V3> V;
        std::vector<Edge> E, E2;
        snap(cs, tol, V, E);
        tsplit(V, E, tol, E2);
        auto lv = loops_of(static_cast<int>(V.size()), E2);
        if (lv.empty()) continue;
        int base = out.loops_raw;
        out.loops_raw += static_cast<int>(lv.size());
        for (int li = 0; li < static_cast<int>(lv.size()); ++li) {
            const auto& L = lv[li];
            if (L.size() < 3) continue;
            LoopOut lo;
            for (int v : L) lo.pts.push_back(V[v]);
            sn_gn(lo.pts, lo.Sn, lo.Gn);
            if (!xc0_of(lo.Sn, lo.Gn, lo.xc, lo.nA, lo.area)) continue;
            if (lo.area < area_eps) continue;
            V3 dir = pB.pos - pA.pos;
            if (dot(lo.nA, dir) < 0.0) lo.nA = lo.nA * -1.0;
            lo.comp = ci;
            lo.raw_index = base + li;
            out.loops.push_back(std::move(lo));
        }
    }
}

// ---- simulation: main.cpp:1512-1529, 1531-1689, 1887-2505 ----
struct PInit {
    std::string stl;
    V3 pos, vel, omega;
    Q rot;
    double scale = 0.0, density = 1.0, young = 1e7, poisson = 0.25, mu = 0.5, rest = 0.5;
};
struct WInit {
    std::string stl;
    V3 pos;
    Q rot;
    double scale = 1.0, mu = 0.5, rest = 0.5;
};
struct Config {
    int n = 2, steps = 1;
    double dt = 1e-4, spacing = 2.5;
    V3 g{0.0, -9.81, 0.0};
    bool split = true, center = true;
    std::string stl;
    int output_interval = 1;
    V3 v0, v1;
    int force_model = 0;       // 1: minimal-energy force, 2: the reference's force with its dormant
                               // normal damping (build extension key, ignored by the reference)
    double k_minimal = 0.0;
    std::vector<PInit> parts;
    std::vector<WInit> walls;
};

bool parse_config(const std::string& path, Config& c) {
    std::ifstream in(path);
    if (!in) return false;
    std::string line;
    bool inp = false, inw = false;
    PInit cp;
    WInit cw;
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
            cp = PInit{}; inp = true; continue;
        }
        if (k == "end_particle" || k == "particle_end") { if (inp) { c.parts.push_back(cp); inp = false; } continue; }
        if (k == "wall") {
            if (inp) { c.parts.push_back(cp); inp = false; }
            if (inw) c.walls.push_back(cw);
            cw = WInit{}; inw = true; continue;
        }
        if (k == "end_wall" || k == "wall_end") { if (inw) { c.walls.push_back(cw); inw = false; } continue; }
        if (inp) {
            if (k == "stl") is >> cp.stl;
            else if (k == "pos") is >> cp.pos.x >> cp.pos.y >> cp.pos.z;
            else if (k == "vel") is >> cp.vel.x >> cp.vel.y >> cp.vel.z;
            else if (k == "omega") is >> cp.omega.x >> cp.omega.y >> cp.omega.z;
            else if (k == "quat") is >> cp.rot.w >> cp.rot.x >> cp.rot.y >> cp.rot.z;
            else if (k == "scale") is >> cp.scale;
            else if (k == "density") is >> cp.density;
            else if (k == "young") is >> cp.young;
            else if (k == "poisson") is >> cp.poisson;
            else if (k == "mu") is >> cp.mu;
            else if (k == "restitution") is >> cp.rest;
        } else if (inw) {
            if (k == "stl") is >> cw.stl;
            else if (k == "pos") is >> cw.pos.x >> cw.pos.y >> cw.pos.z;
            else if (k == "quat") is >> cw.rot.w >> cw.rot.x >> cw.rot.y >> cw.rot.z;
            else if (k == "scale") is >> cw.scale;
            else if (k == "mu") is >> cw.mu;
            else if (k == "restitution") is >> cw.rest;
        } else if (k == "stl") is >> c.stl;
        else if (k == "n") is >> c.n;
        else if (k == "steps") is >> c.steps;
        else if (k == "dt") is >> c.dt;
        else if (k == "spacing") is >> c.spacing;
        else if (k == "split_contacts") { int v = 1; is >> v; c.split = (v != 0); }
        else if (k == "gravity") is >> c.g.x >> c.g.y >> c.g.z;
        else if (k == "center_mesh") { int v = 1; is >> v; c.center = (v != 0); }
        else if (k == "force_model") {
            std::string v;
            is >> v;
            c.force_model = v == "minimal" ? 1 : v == "zdem_damped" ? 2 : 0;
        }
        else if (k == "minimal_k") is >> c.k_minimal;
        else if (k == "output_interval") is >> c.output_interval;
        else if (k == "v0") is >> c.v0.x >> c.v0.y >> c.v0.z;
        else if (k == "v1") is >> c.v1.x >> c.v1.y >> c.v1.z;
    }
    if (inp) c.parts.push_back(cp);
    if (inw) c.walls.push_back(cw);
    if (c.output_interval < 1) c.output_interval = 1;
    if (!c.stl.empty()) return true;
    return !c.parts.empty();
}

struct Particle {
    Pose tf;
    V3 vel, omega, L;
    double mass = 1, inv_mass = 1;
    M3 Ib, Ibinv;
    double radius = 0, eqr = 0, young = 0, poisson = 0, mu = 0, rest = 0;
    int mesh = 0;
};
struct Wall {
    Mesh mesh;
    Pose tf;
    double mu = 0.5, rest = 0.5;
    std::vector<V3> tn;
};

M3 inertia_world(const Particle& p) { M3 R = q_to_m(p.tf.rot); return m_mul(m_mul(R, p.Ib), m_T(R)); }
M3 inertia_world_inv(const Particle& p) { M3 R = q_to_m(p.tf.rot); return m_mul(m_mul(R, p.Ibinv), m_T(R)); }
// main.cpp:1494-1510
void integrate(Particle& p, const V3& F, const V3& T, const V3& g, double dt) {
    V3 acc = F * p.inv_mass + g;
    p.vel += acc * dt;
    p.tf.pos += p.vel * dt;
    M3 Iinv = inertia_world_inv(p);
    p.L += T * dt;
    p.omega = m_mulv(Iinv, p.L);
    Q dq{0.0, p.omega.x, p.omega.y, p.omega.z};
    Q qd = qmul(dq, p.tf.rot);
    p.tf.rot.w += 0.5 * qd.w * dt;
    p.tf.rot.x += 0.5 * qd.x * dt;
    p.tf.rot.y += 0.5 * qd.y * dt;
    p.tf.rot.z += 0.5 * qd.z * dt;
    p.tf.rot = qnormalize(p.tf.rot);
}

std::string join_path(const char* base, const std::string& p) {
    if (!base || p.empty() || p[0] == '/') return p;
    std::string b(base);
    if (!b.empty() && b.back() != '/') b.push_back('/');
    return b + p;
}

}  // namespace

struct orc_sim {
    Config cfg;
    std::vector<Mesh> meshes;
    std::vector<Particle> ps;
    std::vector<Wall> walls;
    std::unordered_map<long long, V3> hist, whist;
    std::vector<V3> F, T;
    std::vector<int32_t> counts;
};

namespace {
void sim_step_once(orc_sim& s, int64_t* ctr) {
    auto& ps = s.ps;
    const auto& cfg = s.cfg;
    int N = static_cast<int>(ps.size());
    std::fill(s.F.begin(), s.F.end(), V3{});
    std::fill(s.T.begin(), s.T.end(), V3{});
    std::fill(s.counts.begin(), s.counts.end(), 0);
    int64_t contacts = 0, tri_checks = 0, nseg = 0, ncomp = 0, nloop = 0, npairs = 0, nwall = 0;
    std::unordered_set<long long> active, wactive;
    // main.cpp:2049-2306
    for (int i = 0; i < N; ++i) {
        for (int j = i + 1; j < N; ++j) {
            Particle& a = ps[i];
            Particle& b = ps[j];
            long long key = (static_cast<long long>(i) << 32) | static_cast<unsigned long long>(j);
            V3 dp = b.tf.pos - a.tf.pos;
            double d2 = dot(dp, dp);
            double rs = a.radius + b.radius;
            if (d2 > rs * rs) continue;
            npairs++;
            PairOut po;
            pair_pipeline(s.meshes[a.mesh], a.tf, s.meshes[b.mesh], b.tf, cfg.split, po);
            tri_checks += po.tri_checks;
            if (po.segs.empty()) continue;
            nseg += static_cast<int64_t>(po.segs.size());
            ncomp += po.comps;
            nloop += po.loops_raw;
            for (const auto& lo : po.loops) {
                if (cfg.force_model == 1) {
                    // Minimal-energy force, polyline.py:676-679: Fn = k/2 |Sn|
                    // along nA, at the loop centroid (-3/2 xc0 up to a multiple
                    // of nA).  No damping, no tangential force, no history.
                    V3 xc = lo.xc * -1.5;
                    V3 rA = xc - a.tf.pos, rB = xc - b.tf.pos;
                    V3 f = lo.nA * (0.5 * cfg.k_minimal * lo.area);
                    s.F[i] -= f;
                    s.F[j] += f;
                    s.T[i] += cross(rA, f * -1.0);
                    s.T[j] += cross(rB, f);
                    active.insert(key);
                    s.counts[i] += 1;
                    s.counts[j] += 1;
                    contacts++;
                    continue;
                }
                // main.cpp:2220-2287
                V3 nA = lo.nA;
                double area = lo.area;
                V3 rA = lo.xc - a.tf.pos, rB = lo.xc - b.tf.pos;
                V3 vA = a.vel + cross(a.omega, rA), vB = b.vel + cross(b.omega, rB);
                V3 vrel = vB - vA;
                double vn = dot(vrel, nA);
                double R1 = std::max(a.eqr, 1e-12), R2 = std::max(b.eqr, 1e-12);
                double E1 = std::max(a.young, 1e-12), E2 = std::max(b.young, 1e-12);
                double nu1 = a.poisson, nu2 = b.poisson;
                double kn = (E1 / R1 + E2 / R2) * area;
                double kt = (E1 / (2.0 * (1.0 + nu1) * R1) + E2 / (2.0 * (1.0 + nu2) * R2)) * area;
                double meff = 1.0 / (a.inv_mass + b.inv_mass);
                double e = std::min(a.rest, b.rest);
                e = std::min(0.9999, std::max(1e-6, e));
                double le = std::log(e);
                double cn = 2.0 * le * std::sqrt(kn * meff) / std::sqrt(le * le + 3.141592653589793 * 3.141592653589793);
                double ct = 2.0 * le * std::sqrt(kt * meff) / std::sqrt(le * le + 3.141592653589793 * 3.141592653589793);
                cn = std::abs(cn);
                ct = std::abs(ct);
                double dV = std::pow(area, 1.5);
                double fn = kn * std::cbrt(dV);
                // force model 2: the normal damping the reference computes and
                // leaves commented out (main.cpp:2244-2248, 2258-2260)
                if (cfg.force_model == 2 && vn < 0.0) fn += -cn * vn;
                if (fn < 0.0) fn = 0.0;
                V3 fnv = nA * fn;
                V3 vt = vrel - nA * vn;
                V3 ds = s.hist[key];
                ds += vt * cfg.dt;
                ds -= nA * dot(ds, nA);
                V3 ft = ds * (-kt) + vt * (-ct);
                double mu = std::min(a.mu, b.mu);
                double ftn = norm(ft), ftm = mu * fn;
                if (ftn > ftm && ftn > 1e-14) {
                    ft = ft * (ftm / ftn);
                    if (kt > 1e-12) ds = ft * (-1.0 / kt);
                }
                s.hist[key] = ds;
                V3 f = fnv + ft;
                s.F[i] -= f;
                s.F[j] += f;
                s.T[i] += cross(rA, f * -1.0);
                s.T[j] += cross(rB, f);
                active.insert(key);
                s.counts[i] += 1;
                s.counts[j] += 1;
                contacts++;
            }
        }
    }
    // main.cpp:2308-2447
    for (int i = 0; i < N; ++i) {
        Particle& p = ps[i];
        const Mesh& pm = s.meshes[p.mesh];
        for (int w = 0; w < static_cast<int>(s.walls.size()); ++w) {
            const Wall& wl = s.walls[w];
            V3 dp = p.tf.pos - wl.tf.pos;
            double d2 = dot(dp, dp);
            double rs = p.radius + wl.mesh.radius;
            if (d2 > rs * rs) continue;
            auto wtris = transform_tris(wl.mesh, wl.tf);
            double tol = tol_of(pm);
            for (std::size_t t = 0; t < wtris.size(); ++t) {
                const Tri& wt = wtris[t];
                V3 wn = wl.tn[t];
                V3 wmn{std::min({wt[0].x, wt[1].x, wt[2].x}), std::min({wt[0].y, wt[1].y, wt[2].y}), std::min({wt[0].z, wt[1].z, wt[2].z})};
                V3 wmx{std::max({wt[0].x, wt[1].x, wt[2].x}), std::max({wt[0].y, wt[1].y, wt[2].y}), std::max({wt[0].z, wt[1].z, wt[2].z})};
                V3 pmn = p.tf.pos - V3{p.radius, p.radius, p.radius};
                V3 pmx = p.tf.pos + V3{p.radius, p.radius, p.radius};
                if (pmx.x < wmn.x || pmn.x > wmx.x || pmx.y < wmn.y || pmn.y > wmx.y || pmx.z < wmn.z || pmn.z > wmx.z) continue;
                double area;
                V3 xc, cn;
                if (!wall_contact(pm, p.tf, wt, wn, tol, area, xc, cn)) continue;
                V3 tp = p.tf.pos - xc;
                if (dot(cn, tp) < 0.0) cn = cn * -1.0;
                V3 rP = xc - p.tf.pos;
                V3 vrel = p.vel + cross(p.omega, rP);
                double vn = dot(vrel, cn);
                double Rp = std::max(p.eqr, 1e-12), Ep = std::max(p.young, 1e-12), nu = p.poisson;
                double kn = (Ep / Rp) * area;
                double kt = (Ep / (2.0 * (1.0 + nu) * Rp)) * area;
                double meff = p.mass;
                double e = p.rest;
                e = std::min(0.9999, std::max(1e-6, e));
                double le = std::log(e);
                double cnn = 2.0 * le * std::sqrt(kn * meff) / std::sqrt(le * le + 3.141592653589793 * 3.141592653589793);
                double ct = 2.0 * le * std::sqrt(kt * meff) / std::sqrt(le * le + 3.141592653589793 * 3.141592653589793);
                cnn = std::abs(cnn);
                ct = std::abs(ct);
                double dV = std::pow(area, 1.5);
                double fe = kn * std::cbrt(dV);
                double fd = -cnn * vn;
                double fn = fe + fd;
                if (fn < 0.0) fn = 0.0;
                V3 fnv = cn * fn;
                V3 vt = vrel - cn * vn;
                long long wk = (static_cast<long long>(i) << 16) | static_cast<long long>(w);
                V3 ds = s.whist[wk];
                ds += vt * cfg.dt;
                ds -= cn * dot(ds, cn);
                V3 ft = ds * (-kt) + vt * (-ct);
                double mu = std::min(p.mu, wl.mu);
                double ftn = norm(ft), ftm = mu * fn;
                if (ftn > ftm && ftn > 1e-14) {
                    ft = ft * (ftm / ftn);
                    if (kt > 1e-12) ds = ft * (-1.0 / kt);
                }
                s.whist[wk] = ds;
                V3 f = fnv + ft;
                s.F[i] += f;
                s.T[i] += cross(rP, f);
                wactive.insert(wk);
                s.counts[i] += 1;
                contacts++;
                nwall++;
            }
        }
    }
    for (int i = 0; i < N; ++i) integrate(ps[i], s.F[i], s.T[i], cfg.g, cfg.dt);
    for (auto it = s.hist.begin(); it != s.hist.end();) it = active.count(it->first) ? std::next(it) : s.hist.erase(it);
    for (auto it = s.whist.begin(); it != s.whist.end();) it = wactive.count(it->first) ? std::next(it) : s.whist.erase(it);
    if (ctr) {
        ctr[0] = contacts; ctr[1] = tri_checks; ctr[2] = nseg; ctr[3] = ncomp; ctr[4] = nloop;
        ctr[5] = npairs; ctr[6] = nwall;
    }
}

void to_tris(const double* t, int T, std::vector<Tri>& out) {
    out.resize(T);
    for (int i = 0; i < T; ++i)
        out[i] = Tri{V3{t[9 * i + 0], t[9 * i + 1], t[9 * i + 2]}, V3{t[9 * i + 3], t[9 * i + 4], t[9 * i + 5]},
                     V3{t[9 * i + 6], t[9 * i + 7], t[9 * i + 8]}};
}
inline void put3(double* o, const V3& v) { o[0] = v.x; o[1] = v.y; o[2] = v.z; }
}  // namespace

extern "C" {

int orc_load_stl(const char* path, double* tris, int cap) {
    std::vector<Tri> t;
    if (!load_stl(path, t)) return -1;
    int T = static_cast<int>(t.size());
    if (tris && T <= cap)
        for (int i = 0; i < T; ++i)
            for (int k = 0; k < 3; ++k) put3(tris + 9 * i + 3 * k, t[i][k]);
    return T;
}

int orc_build_mesh(const double* tris, int T, int center, double* out_tris, double* props) {
    std::vector<Tri> t;
    to_tris(tris, T, t);
    Mesh m = build_mesh(t, center != 0);
    if (out_tris)
        for (int i = 0; i < T; ++i)
            for (int k = 0; k < 3; ++k) put3(out_tris + 9 * i + 3 * k, m.tris[i][k]);
    if (props) {
        props[0] = m.volume; put3(props + 1, m.center);
        props[4] = m.radius; props[5] = m.mean_edge; props[6] = m.bbox_diag;
        for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) props[7 + 3 * i + j] = m.inertia_unit.m[i][j];
    }
    return static_cast<int>(m.verts.size());
}

int orc_pair_contact(const double* trisA, int TA, const double* trisB, int TB,
                     const double* posA, const double* quatA, const double* posB, const double* quatB,
                     double mean_edge_a, double bbox_diag_a, int split_contacts, double k_stiff,
                     double* seg, int seg_cap, double* loop_rec, int loop_cap, double* loop_vtx, int vtx_cap,
                     orc_pair_counts* counts) {
    Mesh mA, mB;
    to_tris(trisA, TA, mA.tris);
    to_tris(trisB, TB, mB.tris);
    mA.mean_edge = mean_edge_a;
    mA.bbox_diag = bbox_diag_a;
    Pose pA{V3{posA[0], posA[1], posA[2]}, Q{quatA[0], quatA[1], quatA[2], quatA[3]}};
    Pose pB{V3{posB[0], posB[1], posB[2]}, Q{quatB[0], quatB[1], quatB[2], quatB[3]}};
    PairOut po;
    pair_pipeline(mA, pA, mB, pB, split_contacts != 0, po);
    orc_pair_counts c{};
    c.tri_checks = po.tri_checks;
    c.segments = static_cast<int32_t>(po.segs.size());
    c.comps = po.comps;
    c.loops_raw = po.loops_raw;
    c.loops = static_cast<int32_t>(po.loops.size());
    int nv = 0;
    for (const auto& lo : po.loops) nv += static_cast<int>(lo.pts.size());
    c.loop_verts = nv;
    if ((seg && c.segments > seg_cap) || (loop_rec && c.loops > loop_cap) || (loop_vtx && nv > vtx_cap)) c.overflow = 1;
    if (seg && !c.overflow)
        for (int s = 0; s < c.segments; ++s) { put3(seg + 6 * s, po.segs[s].first); put3(seg + 6 * s + 3, po.segs[s].second); }
    if (loop_rec && !c.overflow) {
        for (int l = 0; l < c.loops; ++l) {
            const auto& lo = po.loops[l];
            double* r = loop_rec + 28 * l;
            put3(r + 0, lo.Sn); put3(r + 3, lo.Gn); put3(r + 6, lo.nA); put3(r + 9, lo.xc);
            r[12] = lo.area;
            // polyline.py:676-679 minimal-energy force, direction as main.cpp:2262/2281-2284
            double Fn = 0.5 * k_stiff * lo.area;
            r[13] = Fn;
            V3 f = lo.nA * Fn;
            put3(r + 14, f);
            V3 rA = lo.xc - pA.pos, rB = lo.xc - pB.pos;
            put3(r + 17, cross(rA, f * -1.0));
            put3(r + 20, cross(rB, f));
            r[23] = lo.comp;
            r[24] = static_cast<double>(lo.pts.size());
            r[25] = lo.raw_index;
            r[26] = 0.0; r[27] = 0.0;
        }
    }
    if (loop_vtx && !c.overflow) {
        int k = 0;
        for (const auto& lo : po.loops) for (const auto& p : lo.pts) { put3(loop_vtx + 3 * k, p); ++k; }
    }
    if (counts) *counts = c;
    return c.overflow ? 1 : 0;
}

orc_sim* orc_sim_from_config(const char* path, const char* base_dir) {
    Config cfg;
    if (!parse_config(path, cfg)) return nullptr;
    auto* s = new orc_sim();
    s->cfg = cfg;
    std::unordered_map<std::string, std::vector<Tri>> tcache;
    std::unordered_map<std::string, int> mcache;
    auto load_for = [&](const std::string& p, double scale) -> int {
        std::string key = p + "|" + std::to_string(scale);
        auto it = mcache.find(key);
        if (it != mcache.end()) return it->second;
        auto jt = tcache.find(p);
        if (jt == tcache.end()) {
            std::vector<Tri> t;
            if (!load_stl(join_path(base_dir, p), t)) return -1;
            tcache[p] = t;
            jt = tcache.find(p);
        }
        std::vector<Tri> t = jt->second;
        for (auto& tr : t) { tr[0] *= scale; tr[1] *= scale; tr[2] *= scale; }
        s->meshes.push_back(build_mesh(t, cfg.center));
        int idx = static_cast<int>(s->meshes.size()) - 1;
        mcache[key] = idx;
        return idx;
    };
    std::vector<PInit> inits;
    if (!cfg.parts.empty()) inits = cfg.parts;
    else {
        inits.resize(cfg.n);
        for (int i = 0; i < cfg.n; ++i) {
            inits[i].stl = cfg.stl;
            inits[i].pos = V3{cfg.spacing * i, 0.0, 0.0};
            inits[i].vel = (i == 0) ? cfg.v0 : (i == 1 ? cfg.v1 : V3{});
            inits[i].scale = 1.0; inits[i].density = 1.0; inits[i].young = 1e7;
            inits[i].poisson = 0.25; inits[i].mu = 0.5; inits[i].rest = 0.5;
        }
    }
    // main.cpp:1936-1987
    for (const auto& in : inits) {
        std::string sp = in.stl.empty() ? cfg.stl : in.stl;
        if (sp.empty()) { delete s; return nullptr; }
        double sc = in.scale > 0.0 ? in.scale : 1.0;
        int mi = load_for(sp, sc);
        if (mi < 0) { delete s; return nullptr; }
        const Mesh& m = s->meshes[mi];
        if (m.verts.empty()) { delete s; return nullptr; }
        Particle p;
        p.tf.pos = in.pos; p.tf.rot = in.rot; p.vel = in.vel; p.omega = in.omega;
        double den = in.density > 0.0 ? in.density : 1.0;
        if (m.volume > 0.0) {
            p.mass = den * m.volume;
            p.inv_mass = 1.0 / p.mass;
            p.Ib = m_scale(m.inertia_unit, den);
            p.Ibinv = m_inv(p.Ib);
        } else {
            p.mass = den;
            p.inv_mass = 1.0 / p.mass;
            double I = 0.4 * p.mass * m.radius * m.radius;
            M3 Ib = m_identity();
            Ib.m[0][0] = I; Ib.m[1][1] = I; Ib.m[2][2] = I;
            p.Ib = Ib;
            p.Ibinv = m_inv(Ib);
        }
        p.radius = m.radius;
        p.eqr = m.volume > 0.0 ? std::cbrt(3.0 * m.volume / (4.0 * 3.141592653589793)) : m.radius;
        p.young = in.young; p.poisson = in.poisson; p.mu = in.mu; p.rest = in.rest;
        p.mesh = mi;
        p.L = m_mulv(inertia_world(p), p.omega);
        s->ps.push_back(p);
    }
    // main.cpp:1990-2022
    for (const auto& wi : cfg.walls) {
        if (wi.stl.empty()) continue;
        std::vector<Tri> wt;
        if (!load_stl(join_path(base_dir, wi.stl), wt)) continue;
        double sc = wi.scale > 0.0 ? wi.scale : 1.0;
        for (auto& t : wt) { t[0] = t[0] * sc; t[1] = t[1] * sc; t[2] = t[2] * sc; }
        Wall w;
        w.mesh = build_mesh(wt, cfg.center);
        w.tf.pos = wi.pos; w.tf.rot = wi.rot; w.mu = wi.mu; w.rest = wi.rest;
        auto wt2 = transform_tris(w.mesh, w.tf);
        for (const auto& t : wt2) w.tn.push_back(tri_normal(t));
        s->walls.push_back(std::move(w));
    }
    s->F.assign(s->ps.size(), V3{});
    s->T.assign(s->ps.size(), V3{});
    s->counts.assign(s->ps.size(), 0);
    return s;
}

void orc_sim_destroy(orc_sim* s) { delete s; }
int orc_sim_num_particles(const orc_sim* s) { return static_cast<int>(s->ps.size()); }
double orc_sim_dt(const orc_sim* s) { return s->cfg.dt; }
void orc_sim_set_force_model(orc_sim* s, int model, double k) {
    s->cfg.force_model = model;
    s->cfg.k_minimal = k;
}

void orc_sim_get_state(const orc_sim* s, double* o) {
    for (std::size_t i = 0; i < s->ps.size(); ++i) {
        const Particle& p = s->ps[i];
        double* r = o + 16 * i;
        put3(r, p.tf.pos);
        r[3] = p.tf.rot.w; r[4] = p.tf.rot.x; r[5] = p.tf.rot.y; r[6] = p.tf.rot.z;
        put3(r + 7, p.vel); put3(r + 10, p.omega); put3(r + 13, p.L);
    }
}
void orc_sim_set_state(orc_sim* s, const double* in) {
    for (std::size_t i = 0; i < s->ps.size(); ++i) {
        Particle& p = s->ps[i];
        const double* r = in + 16 * i;
        p.tf.pos = V3{r[0], r[1], r[2]};
        p.tf.rot = Q{r[3], r[4], r[5], r[6]};
        p.vel = V3{r[7], r[8], r[9]};
        p.omega = V3{r[10], r[11], r[12]};
        p.L = V3{r[13], r[14], r[15]};
    }
}
void orc_sim_get_forces(const orc_sim* s, double* F, double* T, int32_t* counts) {
    for (std::size_t i = 0; i < s->ps.size(); ++i) {
        if (F) put3(F + 3 * i, s->F[i]);
        if (T) put3(T + 3 * i, s->T[i]);
        if (counts) counts[i] = s->counts[i];
    }
}
int orc_sim_step(orc_sim* s, int n, int64_t* ctr) {
    for (int k = 0; k < n; ++k) sim_step_once(*s, ctr);
    return 0;
}

int orc_tri_tri(const double* a, const double* b, double* out) {
    std::vector<Tri> A, B;
    to_tris(a, 1, A);
    to_tris(b, 1, B);
    V3 nA, nB;
    double dA = 0, dB = 0;
    if (!plane_from_tri(A[0], nA, dA)) { nA = V3{}; dA = 0; }
    if (!plane_from_tri(B[0], nB, dB)) { nB = V3{}; dB = 0; }
    if (norm2(nA) < 1e-30 || norm2(nB) < 1e-30) return 0;
    V3 p0, p1;
    if (!tri_tri(A[0], nA, dA, B[0], nB, dB, p0, p1)) return 0;
    put3(out, p0); put3(out + 3, p1);
    return 1;
}
void orc_sn_gn(const double* pts, int m, double* Sn, double* Gn) {
    std::vector<V3> L(m);
    for (int i = 0; i < m; ++i) L[i] = V3{pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]};
    V3 a, b;
    sn_gn(L, a, b);
    put3(Sn, a); put3(Gn, b);
}
int orc_xc0(const double* Sn, const double* Gn, double* xc, double* nA, double* area) {
    V3 x, n;
    bool ok = xc0_of(V3{Sn[0], Sn[1], Sn[2]}, V3{Gn[0], Gn[1], Gn[2]}, x, n, *area);
    put3(xc, x); put3(nA, n);
    return ok ? 1 : 0;
}
void orc_loop_start_order(const int32_t* deg, int n, int32_t* order) {
    std::vector<int> d(deg, deg + n), o;
    start_order(d, o);
    for (int i = 0; i < n; ++i) order[i] = o[i];
}
int orc_sh_clip(const double* subj, int ns, const double* clip, int nc, double* out, int cap) {
    std::vector<V2> s(ns), c(nc);
    for (int i = 0; i < ns; ++i) s[i] = V2{subj[2 * i], subj[2 * i + 1]};
    for (int i = 0; i < nc; ++i) c[i] = V2{clip[2 * i], clip[2 * i + 1]};
    auto o = sh_clip(s, c);
    int n = static_cast<int>(o.size());
    for (int i = 0; i < n && i < cap; ++i) { out[2 * i] = o[i].x; out[2 * i + 1] = o[i].y; }
    return n;
}
int orc_wall_contact(const double* tris_body, int T, const double* pos, const double* quat,
                     const double* wall_tri, const double* wall_n, double tol,
                     double* area, double* xc, double* n) {
    Mesh m;
    to_tris(tris_body, T, m.tris);
    Pose p{V3{pos[0], pos[1], pos[2]}, Q{quat[0], quat[1], quat[2], quat[3]}};
    std::vector<Tri> w;
    to_tris(wall_tri, 1, w);
    V3 x, nn;
    if (!wall_contact(m, p, w[0], V3{wall_n[0], wall_n[1], wall_n[2]}, tol, *area, x, nn)) return 0;
    put3(xc, x); put3(n, nn);
    return 1;
}
void orc_integrate(double* st, const double* Ibinv, double inv_mass, const double* F, const double* T,
                   const double* g, double dt) {
    Particle p;
    p.tf.pos = V3{st[0], st[1], st[2]};
    p.tf.rot = Q{st[3], st[4], st[5], st[6]};
    p.vel = V3{st[7], st[8], st[9]};
    p.omega = V3{st[10], st[11], st[12]};
    p.L = V3{st[13], st[14], st[15]};
    p.inv_mass = inv_mass;
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) p.Ibinv.m[i][j] = Ibinv[3 * i + j];
    integrate(p, V3{F[0], F[1], F[2]}, V3{T[0], T[1], T[2]}, V3{g[0], g[1], g[2]}, dt);
    put3(st, p.tf.pos);
    st[3] = p.tf.rot.w; st[4] = p.tf.rot.x; st[5] = p.tf.rot.y; st[6] = p.tf.rot.z;
    put3(st + 7, p.vel); put3(st + 10, p.omega); put3(st + 13, p.L);
}

}  // extern "C"
