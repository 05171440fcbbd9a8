// zdx_geom.h — FP64 geometry of the Route-B contact step, __host__ __device__.
//
// Every expression keeps the operation order of the reference
// (zhl-handsome/ZEDM-X src/main.cpp, cited per function) and the library is
// compiled with -ffp-contract=off, so add/sub/mul/div/sqrt results are
// bit-identical to the reference's x86-64 build.  The per-pair stages after
// the narrow phase (components, snap, T-junction split, loop tracing, Sn/Gn
// reduction) are written as one routine that a single GPU thread runs per
// pair over strided per-thread scratch; the same source compiles on the host
// for CPU-side checks.
#pragma once
#include <math.h>
#include <stdint.h>

#ifndef ZDX_HD
#define ZDX_HD __host__ __device__
#endif
#include "zdx_sort.h"

namespace zdx {

// ------------------------------------------------------------ primitives
struct V3 { double x, y, z; };
struct Qt { double w, x, y, z; };

ZDX_HD inline V3 v3(double x, double y, double z) { V3 r; r.x = x; r.y = y; r.z = z; return r; }
ZDX_HD inline V3 add(const V3& a, const V3& b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
ZDX_HD inline V3 sub(const V3& a, const V3& b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
ZDX_HD inline V3 mul(const V3& a, double s) { return v3(a.x * s, a.y * s, a.z * s); }
ZDX_HD inline V3 dvd(const V3& a, double s) { return v3(a.x / s, a.y / s, a.z / s); }
// main.cpp:37-63
ZDX_HD inline double dot(const V3& a, const V3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
ZDX_HD inline V3 cross(const V3& a, const V3& b) {
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
ZDX_HD inline double norm2(const V3& v) { return dot(v, v); }
ZDX_HD inline double norm(const V3& v) { return sqrt(dot(v, v)); }
ZDX_HD inline V3 normalize(const V3& v) {
    double n = norm(v);
    if (n < 1e-14) return v3(0.0, 0.0, 0.0);
    return dvd(v, n);
}
// std::min / std::max semantics (return the first argument on ties)
ZDX_HD inline double dmin(double a, double b) { return (b < a) ? b : a; }
ZDX_HD inline double dmax(double a, double b) { return (a < b) ? b : a; }
// std::min({a,b,c}) / std::max({a,b,c}): first extreme element
ZDX_HD inline double min3(double a, double b, double c) { double m = a; if (b < m) m = b; if (c < m) m = c; return m; }
ZDX_HD inline double max3(double a, double b, double c) { double m = a; if (m < b) m = b; if (m < c) m = c; return m; }

// main.cpp:72-98
ZDX_HD inline Qt qmul(const Qt& a, const Qt& b) {
    Qt r;
    r.w = a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z;
    r.x = a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y;
    r.y = a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x;
    r.z = a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w;
    return r;
}
ZDX_HD inline Qt qnormalize(const Qt& q) {
    double n = sqrt(q.w * q.w + q.x * q.x + q.y * q.y + q.z * q.z);
    Qt r;
    if (n < 1e-14) { r.w = 1.0; r.x = 0.0; r.y = 0.0; r.z = 0.0; return r; }
    r.w = q.w / n; r.x = q.x / n; r.y = q.y / n; r.z = q.z / n;
    return r;
}
// quat_rotate: q normalised on every call, then qn * (0,v) * conj(qn)
ZDX_HD inline V3 qrotate(const Qt& q, const V3& v) {
    Qt p; p.w = 0.0; p.x = v.x; p.y = v.y; p.z = v.z;
    Qt qn = qnormalize(q);
    Qt c; c.w = qn.w; c.x = -qn.x; c.y = -qn.y; c.z = -qn.z;
    Qt r = qmul(qmul(qn, p), c);
    return v3(r.x, r.y, r.z);
}

struct M3 { double m[3][3]; };
// quat_to_mat3 main.cpp:150-164
ZDX_HD inline M3 q_to_m(const Qt& q) {
    Qt n = qnormalize(q);
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
// mat3_mul main.cpp:174-185 (accumulate from 0.0 in k order)
ZDX_HD inline M3 m_mul(const M3& a, const M3& b) {
    M3 r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double s = 0.0;
            for (int k = 0; k < 3; ++k) s += a.m[i][k] * b.m[k][j];
            r.m[i][j] = s;
        }
    return r;
}
ZDX_HD inline M3 m_T(const M3& a) {
    M3 r;
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) r.m[i][j] = a.m[j][i];
    return r;
}
ZDX_HD inline V3 m_mulv(const M3& m, const V3& v) {
    return v3(m.m[0][0] * v.x + m.m[0][1] * v.y + m.m[0][2] * v.z,
              m.m[1][0] * v.x + m.m[1][1] * v.y + m.m[1][2] * v.z,
              m.m[2][0] * v.x + m.m[2][1] * v.y + m.m[2][2] * v.z);
}

// ------------------------------------------------------------ particle state
// Rigid-body state of one particle, packed as 16 doubles:
// pos(3) quat(4: w x y z) vel(3) omega(3) L(3).
struct Body {
    V3 pos; Qt q; V3 vel, omega, L;
};
ZDX_HD inline Body load_body(const double* s) {
    Body b;
    b.pos = v3(s[0], s[1], s[2]);
    b.q.w = s[3]; b.q.x = s[4]; b.q.y = s[5]; b.q.z = s[6];
    b.vel = v3(s[7], s[8], s[9]);
    b.omega = v3(s[10], s[11], s[12]);
    b.L = v3(s[13], s[14], s[15]);
    return b;
}
ZDX_HD inline void store_body(double* s, const Body& b) {
    s[0] = b.pos.x; s[1] = b.pos.y; s[2] = b.pos.z;
    s[3] = b.q.w; s[4] = b.q.x; s[5] = b.q.y; s[6] = b.q.z;
    s[7] = b.vel.x; s[8] = b.vel.y; s[9] = b.vel.z;
    s[10] = b.omega.x; s[11] = b.omega.y; s[12] = b.omega.z;
    s[13] = b.L.x; s[14] = b.L.y; s[15] = b.L.z;
}

// integrate_particle main.cpp:1494-1510 with inertia_world_inv 1488-1492
ZDX_HD inline void integrate(Body& p, const M3& Ibinv, double inv_mass, const V3& F, const V3& T,
                             const V3& g, double dt) {
    V3 acc = add(mul(F, inv_mass), g);
    p.vel = add(p.vel, mul(acc, dt));
    p.pos = add(p.pos, mul(p.vel, dt));
    M3 R = q_to_m(p.q);
    M3 Iinv = m_mul(m_mul(R, Ibinv), m_T(R));
    p.L = add(p.L, mul(T, dt));
    p.omega = m_mulv(Iinv, p.L);
    Qt dq; dq.w = 0.0; dq.x = p.omega.x; dq.y = p.omega.y; dq.z = p.omega.z;
    Qt qd = qmul(dq, p.q);
    p.q.w += 0.5 * qd.w * dt;
    p.q.x += 0.5 * qd.x * dt;
    p.q.y += 0.5 * qd.y * dt;
    p.q.z += 0.5 * qd.z * dt;
    p.q = qnormalize(p.q);
}

// ------------------------------------------------------------ triangles
// plane_from_tri main.cpp:518-527; returns false (n = 0, d = 0) when degenerate
ZDX_HD inline bool plane_of(const V3& a, const V3& b, const V3& c, V3& n, double& d) {
    n = cross(sub(b, a), sub(c, a));
    double ln = norm(n);
    if (ln < 1e-30) { n = v3(0.0, 0.0, 0.0); d = 0.0; return false; }
    n = dvd(n, ln);
    d = -dot(n, a);
    return true;
}
// point_in_tri main.cpp:529-543
ZDX_HD inline bool point_in_tri(const V3& p, const V3* t, const V3& n) {
    V3 a = t[0], b = t[1], c = t[2];
    V3 ab = sub(b, a), bc = sub(c, b), ca = sub(a, c);
    V3 ap = sub(p, a), bp = sub(p, b), cp = sub(p, c);
    double c1 = dot(cross(ab, ap), n);
    double c2 = dot(cross(bc, bp), n);
    double c3 = dot(cross(ca, cp), n);
    return (c1 >= -1e-10) && (c2 >= -1e-10) && (c3 >= -1e-10);
}
// seg_plane_intersection main.cpp:545-559
ZDX_HD inline bool seg_plane(const V3& p0, const V3& p1, const V3& n, double d, V3& out) {
    V3 u = sub(p1, p0);
    double den = dot(n, u);
    double num = -(dot(n, p0) + d);
    if (fabs(den) < 1e-30) return false;
    double t = num / den;
    if (t < -1e-10 || t > 1.0 + 1e-10) return false;
    t = dmin(1.0, dmax(0.0, t));
    out = add(p0, mul(u, t));
    return true;
}
// tri_tri_intersection_segment main.cpp:579-631 (unique_points 561-577 inlined
// on a fixed array of at most 6 candidates)
ZDX_HD inline bool tri_tri(const V3* A, const V3& nA, double dA, const V3* B, const V3& nB, double dB,
                           V3& o0, V3& o1) {
    V3 tau = cross(nA, nB);
    double lt = norm(tau);
    if (lt < 1e-20) return false;
    tau = dvd(tau, lt);
    V3 raw[6];
    int nr = 0;
    for (int i = 0; i < 3; ++i) {
        V3 p;
        if (seg_plane(A[i], A[(i + 1) % 3], nB, dB, p) && point_in_tri(p, B, nB)) raw[nr++] = p;
    }
    for (int i = 0; i < 3; ++i) {
        V3 p;
        if (seg_plane(B[i], B[(i + 1) % 3], nA, dA, p) && point_in_tri(p, A, nA)) raw[nr++] = p;
    }
    V3 c[6];
    int nc = 0;
    const double t2 = 1e-9 * 1e-9;
    for (int i = 0; i < nr; ++i) {
        bool keep = true;
        for (int j = 0; j < nc; ++j)
            if (norm2(sub(raw[i], c[j])) <= t2) { keep = false; break; }
        if (keep) c[nc++] = raw[i];
    }
    if (nc < 2) return false;
    double mn = INFINITY, mx = -INFINITY;
    int i0 = -1, i1 = -1;
    for (int i = 0; i < nc; ++i) {
        double t = dot(c[i], tau);
        if (t < mn) { mn = t; i0 = i; }
        if (t > mx) { mx = t; i1 = i; }
    }
    if (i0 == i1) return false;
    V3 p0 = c[i0], p1 = c[i1];
    if (norm2(sub(p1, p0)) < 1e-18) return false;
    if (dot(sub(p1, p0), tau) < 0.0) { V3 t = p0; p0 = p1; p1 = t; }
    o0 = p0; o1 = p1;
    return true;
}

// ------------------------------------------------------------ strided scratch
// Per-thread scratch arrays are interleaved across the threads of a launch
// (element k of thread t at base[k * stride + t]) so that lanes of a wave that
// touch the same k hit consecutive addresses.
template <class T>
struct SArr {
    using value_type = T;
    T* p;
    int64_t s;
    ZDX_HD T get(int i) const { return p[(int64_t)i * s]; }
    ZDX_HD void set(int i, T v) { p[(int64_t)i * s] = v; }
    ZDX_HD T& operator[](int i) const { return p[(int64_t)i * s]; }
    ZDX_HD SArr<T> at(int64_t off) const { SArr<T> r; r.p = p + off * s; r.s = s; return r; }
};

struct I3 { int x, y, z; };
// cell_index main.cpp:649-655
ZDX_HD inline I3 cell_of(const V3& p, double h) {
    I3 c;
    c.x = (int)floor(p.x / h);
    c.y = (int)floor(p.y / h);
    c.z = (int)floor(p.z / h);
    return c;
}
ZDX_HD inline bool near1(int ax, int ay, int az, int bx, int by, int bz) {
    int dx = bx - ax, dy = by - ay, dz = bz - az;
    return dx >= -1 && dx <= 1 && dy >= -1 && dy <= 1 && dz >= -1 && dz <= 1;
}

// seg_seg_dist2 main.cpp:761-802
ZDX_HD inline double seg_seg(const V3& p1, const V3& q1, const V3& p2, const V3& q2) {
    V3 d1 = sub(q1, p1), d2 = sub(q2, p2), r = sub(p1, p2);
    double a = dot(d1, d1), e = dot(d2, d2), f = dot(d2, r);
    if (a <= 1e-12 && e <= 1e-12) return dot(sub(p1, p2), sub(p1, p2));
    if (a <= 1e-12) {
        double t = dmin(1.0, dmax(0.0, f / e));
        V3 c = add(p2, mul(d2, t));
        return dot(sub(p1, c), sub(p1, c));
    }
    if (e <= 1e-12) {
        double s = dmin(1.0, dmax(0.0, -dot(d1, r) / a));
        V3 c = add(p1, mul(d1, s));
        return dot(sub(c, p2), sub(c, p2));
    }
    double b = dot(d1, d2), c = dot(d1, r);
    double den = a * e - b * b;
    double s = 0.0;
    if (fabs(den) > 0.0) s = dmin(1.0, dmax(0.0, (b * f - c * e) / den));
    double t = (b * s + f) / e;
    if (t < 0.0) { t = 0.0; s = dmin(1.0, dmax(0.0, -c / a)); }
    else if (t > 1.0) { t = 1.0; s = dmin(1.0, dmax(0.0, (b - c) / a)); }
    V3 cp1 = add(p1, mul(d1, s));
    V3 cp2 = add(p2, mul(d2, t));
    V3 df = sub(cp1, cp2);
    return dot(df, df);
}
// point_segment_dist2_and_t main.cpp:748-759
ZDX_HD inline void pt_seg(const V3& p, const V3& a, const V3& b, double& d2, double& t) {
    V3 ab = sub(b, a);
    double den = dot(ab, ab);
    if (den <= 1e-12) { d2 = norm2(sub(p, a)); t = 0.0; return; }
    double tt = dot(sub(p, a), ab) / den;
    double tc = dmin(1.0, dmax(0.0, tt));
    V3 pr = add(a, mul(ab, tc));
    d2 = norm2(sub(p, pr));
    t = tc;
}

// ------------------------------------------------------------ per-pair topology
// Scratch sizes for a pair with at most S segments (caps checked at run time).
struct TopoCaps {
    int S;       // segments
    int E2;      // edges after T-junction splitting (4*S)
    // sizes of the int / double scratch regions laid out in topology()
    ZDX_HD int n_ints() const {
        const int S2 = 2 * S;
        const int adj = (S * S + 3) / 4;            // S x S candidate bytes
        return 9 * S + 4 * S + 2 * (S + 1) + 3 * S2 + 5 * S2 + 2 * S2 + 3 * S2 + 2 * E2 + 3 * (2 * S2) + 4 * E2 +
               4 * S2 + E2 + E2 + 1 + adj + 8;
    }
    ZDX_HD int n_dbls() const { return 3 * (2 * S) + 2 * (2 * S) + 8; }
};

struct Split { double t; int v; int k; };   // k = reference scan-order key

struct SplitArr {
    using value_type = Split;
    SArr<double> t;
    SArr<int> v;
    SArr<int> k;
    ZDX_HD Split get(int i) const { Split s; s.t = t.get(i); s.v = v.get(i); s.k = k.get(i); return s; }
    ZDX_HD void set(int i, const Split& s) { t.set(i, s.t); v.set(i, s.v); k.set(i, s.k); }
};
struct SplitByT { ZDX_HD bool operator()(const Split& a, const Split& b) const { return a.t < b.t; } };
struct SplitByScan { ZDX_HD bool operator()(const Split& a, const Split& b) const { return a.k < b.k || (a.k == b.k && a.v < b.v); } };

struct CompBySizeDesc {
    SArr<int> size;
    ZDX_HD bool operator()(int a, int b) const { return size.get(a) > size.get(b); }
};
// extract_loops start-vertex comparator main.cpp:977-982
struct StartOrder {
    SArr<int> deg;
    ZDX_HD bool operator()(int i, int j) const {
        int di = deg.get(i), dj = deg.get(j);
        if ((di == 2) != (dj == 2)) return di == 2;
        return di > dj;
    }
};

// Status bits reported per pair.
enum : int { TOPO_OK = 0, TOPO_E2_OVERFLOW = 1, TOPO_LOOP_OVERFLOW = 2, TOPO_VTX_OVERFLOW = 4 };

struct TopoOut {
    int comps;
    int loops_raw;
    int loops;       // accepted
    int status;
};

// Lane policies for topology(): W lanes run the pair's algorithm together.
// Control flow and every serial step are executed redundantly by all lanes
// with identical values (each lane reads back its own identical writes), and
// the O(S^2) scans are split across lanes, their results combined by a ballot
// mask that the lanes then walk in ascending order.  sync() is needed only
// after a lane-parallel phase whose writes other lanes read.
// mark(k) closes the span of phase k (a no-op unless a policy times phases).
struct LaneHost {
    static constexpr int W = 1;
    int lane() const { return 0; }
    uint64_t ballot(bool p) const { return p ? 1ull : 0ull; }
    double shfl(double v, int) const { return v; }
    void sync() const {}
    void mark(int) const {}
};

ZDX_HD inline int ctz64(uint64_t m) {
    int k = 0;
    while (!(m & 1ull)) { m >>= 1; ++k; }
    return k;
}

// Reference scan-order key of candidate j for segment i in
// label_segments_by_contact_greedy (main.cpp:834-857): rep point r of i (p0,
// p1, mid), neighbour cell (dx,dy,dz) lexicographic; smallest match wins.
ZDX_HD inline int cand_key(SArr<int> cell9, int i, int j) {
    int key = 1 << 30;
    for (int r = 0; r < 3; ++r) {
        int ax = cell9[9 * i + 3 * r], ay = cell9[9 * i + 3 * r + 1], az = cell9[9 * i + 3 * r + 2];
        for (int q = 0; q < 3; ++q) {
            int bx = cell9[9 * j + 3 * q], by = cell9[9 * j + 3 * q + 1], bz = cell9[9 * j + 3 * q + 2];
            if (near1(ax, ay, az, bx, by, bz)) {
                int k = r * 27 + (bx - ax + 1) * 9 + (by - ay + 1) * 3 + (bz - az + 1);
                if (k < key) key = k;
            }
        }
    }
    return key;
}

// Per-pair pipeline after the narrow phase (main.cpp:2152-2219):
// components (label_segments_by_contact_greedy 804-898, when split), then per
// component snap_endpoints (680-746), split_t_junctions (900-961),
// extract_loops (963-1025), and per loop accumulate_Sn_Gn_from_polyline
// (1027-1040), contact_point_xc0 (1042-1052), the <3-vertex and area_eps
// filters and the flip of nA toward B (2188-2219).
//
// The quadratic searches run as flat lane-parallel passes whose iterations do
// not depend on each other: the component test of every ordered segment pair
// (popped i, candidate j -- the order the reference evaluates it in), the
// snap-distance test of every point pair, and the T-junction test of every
// (edge, vertex) pair.  The order-dependent walks (DFS, vertex numbering,
// split chains, loop tracing, sums) then run on their results.
//
// seg: S segments, 6 doubles each (p0, p1), in narrow-phase order.
// I / D: int and double scratch of TopoCaps::n_ints() / n_dbls() entries,
// contiguous (stride 1).
// For every accepted loop, sink(loop_index, comp, raw_index, Sn, Gn, nA, xc0,
// area, nverts, vids, V) is called (by every lane) in the reference's order.
template <class WP, class Sink>
ZDX_HD inline TopoOut topology(const WP& wp, const double* seg, int S, double tol, double area_eps, bool split,
                               const V3& dirAB, SArr<int> I, SArr<double> D, const TopoCaps& cap, Sink& sink) {
    TopoOut out; out.comps = 0; out.loops_raw = 0; out.loops = 0; out.status = TOPO_OK;
    if (S <= 0) return out;
    const int W = WP::W;
    const int ln = wp.lane();
    const double h = tol;
    const double t2 = h * h;
    const int S2 = 2 * S;
    // ---- scratch layout
    const int E2c = cap.E2;
    const int HC = 2 * S2;                  // T-junction hit capacity
    int o = 0;
    SArr<int> cell9 = I.at(o); o += 9 * S;
    SArr<int> vis = I.at(o); o += S;
    SArr<int> stk = I.at(o); o += S;
    SArr<int> cord = I.at(o); o += S;       // element order of all comps, concatenated
    SArr<int> cidx = I.at(o); o += S;       // comp permutation (sorted)
    SArr<int> cbeg = I.at(o); o += S + 1;   // comp start in cord
    SArr<int> csz = I.at(o); o += S + 1;    // comp sizes
    SArr<int> pcell = I.at(o); o += 3 * S2;
    SArr<int> par = I.at(o); o += S2;
    SArr<int> rnk = I.at(o); o += S2;
    SArr<int> rootvid = I.at(o); o += S2;
    SArr<int> ptv = I.at(o); o += S2;       // vid of each point
    SArr<int> cnt = I.at(o); o += S2;
    SArr<int> E = I.at(o); o += 2 * S2;     // snapped edges (a, b)
    SArr<int> vcell3 = I.at(o); o += 3 * S2;
    SArr<int> E2 = I.at(o); o += 2 * E2c;   // split edges (a, b)
    SArr<int> spv = I.at(o); o += HC;       // T-junction hits: vertex, scan key, edge
    SArr<int> spk = I.at(o); o += HC;
    SArr<int> spe = I.at(o); o += HC;
    SArr<int> adjn = I.at(o); o += 2 * E2c;
    SArr<int> adje = I.at(o); o += 2 * E2c;
    SArr<int> deg = I.at(o); o += S2;
    SArr<int> aoff = I.at(o); o += S2;
    SArr<int> afill = I.at(o); o += S2;
    SArr<int> sord = I.at(o); o += S2;
    SArr<int> used = I.at(o); o += E2c;
    SArr<int> path = I.at(o); o += E2c + 1;
    uint8_t* adjm = (uint8_t*)(I.p + o);    // S x S component-candidate bytes
    int od = 0;
    SArr<double> Vs = D.at(od); od += 3 * S2;
    SArr<double> spt = D.at(od); od += HC;

    auto P0 = [&](int s) { return v3(seg[6 * s + 0], seg[6 * s + 1], seg[6 * s + 2]); };
    auto P1 = [&](int s) { return v3(seg[6 * s + 3], seg[6 * s + 4], seg[6 * s + 5]); };

    // ---- components
    int ncomp = 0;
    if (split) {
        for (int i = ln; i < S; i += W) {
            V3 p0 = P0(i), p1 = P1(i);
            V3 mid = mul(add(p0, p1), 0.5);
            I3 a = cell_of(p0, h), b = cell_of(p1, h), c = cell_of(mid, h);
            cell9[9 * i + 0] = a.x; cell9[9 * i + 1] = a.y; cell9[9 * i + 2] = a.z;
            cell9[9 * i + 3] = b.x; cell9[9 * i + 4] = b.y; cell9[9 * i + 5] = b.z;
            cell9[9 * i + 6] = c.x; cell9[9 * i + 7] = c.y; cell9[9 * i + 8] = c.z;
            vis[i] = 0;
        }
        wp.sync();
        // candidate test of every ordered pair (i popped, j candidate): in the
        // grid neighbourhood, AABB(+tol) overlap, seg_seg_dist2 <= tol^2
        const int SS = S * S;
        for (int k = ln; k < SS; k += W) {
            const int i = k / S, j = k - (k / S) * S;
            bool acc = false;
            if (j != i && cand_key(cell9, i, j) != (1 << 30)) {
                V3 p1 = P0(i), q1 = P1(i), p2 = P0(j), q2 = P1(j);
                V3 amn = v3(dmin(p1.x, q1.x) - h, dmin(p1.y, q1.y) - h, dmin(p1.z, q1.z) - h);
                V3 amx = v3(dmax(p1.x, q1.x) + h, dmax(p1.y, q1.y) + h, dmax(p1.z, q1.z) + h);
                V3 bmn = v3(dmin(p2.x, q2.x) - h, dmin(p2.y, q2.y) - h, dmin(p2.z, q2.z) - h);
                V3 bmx = v3(dmax(p2.x, q2.x) + h, dmax(p2.y, q2.y) + h, dmax(p2.z, q2.z) + h);
                bool ovl = !(amx.x < bmn.x || amn.x > bmx.x || amx.y < bmn.y || amn.y > bmx.y || amx.z < bmn.z ||
                             amn.z > bmx.z);
                acc = ovl && seg_seg(p1, q1, p2, q2) <= t2;
            }
            adjm[k] = acc ? 1 : 0;
        }
        wp.sync();
        wp.mark(1);
        int nord = 0;
        for (int seed = 0; seed < S; ++seed) {
            if (vis[seed]) continue;
            cbeg[ncomp] = nord;
            int sp = 0;
            stk[sp++] = seed;
            vis[seed] = 1;
            cord[nord++] = seed;
            while (sp > 0) {
                const int i = stk[--sp];
                // accepted candidates of i are pushed ordered by (scan key, id)
                const int first_acc = sp;
                for (int base = 0; base < S; base += W) {
                    const int j = base + ln;
                    const bool acc = j < S && adjm[i * S + j] && !vis[j];
                    uint64_t m = wp.ballot(acc);
                    while (m) {
                        const int jj = base + ctz64(m);
                        m &= m - 1;
                        const int key = cand_key(cell9, i, jj);
                        int pos = sp;
                        while (pos > first_acc && spk[pos - 1] > key) {
                            stk[pos] = stk[pos - 1];
                            spk[pos] = spk[pos - 1];
                            --pos;
                        }
                        stk[pos] = jj;
                        spk[pos] = key;
                        ++sp;
                    }
                }
                for (int k = first_acc; k < sp; ++k) {
                    vis[stk[k]] = 1;
                    cord[nord++] = stk[k];
                }
            }
            csz[ncomp] = nord - cbeg[ncomp];
            ++ncomp;
        }
        for (int c = 0; c < ncomp; ++c) cidx[c] = c;
        CompBySizeDesc cmp; cmp.size = csz;
        ref_sort(cidx, ncomp, cmp);
    } else {
        for (int s = 0; s < S; ++s) cord[s] = s;
        cbeg[0] = 0; csz[0] = S; cidx[0] = 0;
        ncomp = 1;
    }
    out.comps = ncomp;
    wp.sync();
    wp.mark(2);

    // ---- per component
    for (int ci = 0; ci < ncomp; ++ci) {
        const int c = cidx[ci];
        const int base = cbeg[c];
        const int ns = csz[c];
        const int n = 2 * ns;
        auto PT = [&](int k) { int s = cord[base + (k >> 1)]; return (k & 1) ? P1(s) : P0(s); };
        // snap_endpoints: union-find over points within tol in +-1 cells; the
        // partition does not depend on the order of the unions
        for (int k = ln; k < n; k += W) {
            I3 cc = cell_of(PT(k), h);
            pcell[3 * k] = cc.x; pcell[3 * k + 1] = cc.y; pcell[3 * k + 2] = cc.z;
            par[k] = k; rnk[k] = 0; rootvid[k] = -1;
        }
        wp.sync();
        auto find = [&](int x) {
            int r = x;
            while (par[r] != r) r = par[r];
            while (par[x] != r) { int nx = par[x]; par[x] = r; x = nx; }
            return r;
        };
        wp.mark(3);
        const int npairs = n * (n - 1) / 2;
        for (int kb = 0; kb < npairs; kb += W) {
            const int k = kb + ln;
            bool hit = false;
            int pi = 0, pj = 0;
            if (k < npairs) {
                // k -> (i, j), j < i, k = i (i - 1) / 2 + j
                int i = (int)((1.0 + sqrt(1.0 + 8.0 * (double)k)) * 0.5);
                while (i * (i - 1) / 2 > k) --i;
                while ((i + 1) * i / 2 <= k) ++i;
                pi = i;
                pj = k - i * (i - 1) / 2;
                hit = near1(pcell[3 * pi], pcell[3 * pi + 1], pcell[3 * pi + 2], pcell[3 * pj], pcell[3 * pj + 1],
                            pcell[3 * pj + 2]) &&
                      norm2(sub(PT(pi), PT(pj))) <= h * h;
            }
            uint64_t m = wp.ballot(hit);
            while (m) {
                const int src = ctz64(m);
                m &= m - 1;
                const int kk = kb + src;
                int i = (int)((1.0 + sqrt(1.0 + 8.0 * (double)kk)) * 0.5);
                while (i * (i - 1) / 2 > kk) --i;
                while ((i + 1) * i / 2 <= kk) ++i;
                const int j = kk - i * (i - 1) / 2;
                int ra = find(i), rb = find(j);
                if (ra != rb) {
                    if (rnk[ra] < rnk[rb]) par[ra] = rb;
                    else if (rnk[ra] > rnk[rb]) par[rb] = ra;
                    else { par[rb] = ra; rnk[ra] += 1; }
                }
            }
            (void)pi; (void)pj;
        }
        wp.mark(4);
        int nv = 0;
        for (int i = 0; i < n; ++i) {
            int r = find(i);
            V3 p = PT(i);
            if (rootvid[r] < 0) {
                rootvid[r] = nv;
                Vs[3 * nv] = p.x; Vs[3 * nv + 1] = p.y; Vs[3 * nv + 2] = p.z;
                cnt[nv] = 1;
                ++nv;
            } else {
                int v = rootvid[r];
                Vs[3 * v] += p.x; Vs[3 * v + 1] += p.y; Vs[3 * v + 2] += p.z;
                cnt[v] += 1;
            }
            ptv[i] = rootvid[r];
        }
        for (int v = 0; v < nv; ++v) {
            double cn = (double)cnt[v];
            Vs[3 * v] = Vs[3 * v] / cn; Vs[3 * v + 1] = Vs[3 * v + 1] / cn; Vs[3 * v + 2] = Vs[3 * v + 2] / cn;
        }
        auto VV = [&](int v) { return v3(Vs[3 * v], Vs[3 * v + 1], Vs[3 * v + 2]); };
        int ne = 0;
        for (int s = 0; s < ns; ++s) {
            int a = ptv[2 * s], b = ptv[2 * s + 1];
            if (a != b) { E[2 * ne] = a; E[2 * ne + 1] = b; ++ne; }
        }
        // split_t_junctions: candidate test of every (edge, vertex) pair, hits
        // collected edge-major in vertex order
        for (int v = 0; v < nv; ++v) {
            I3 cc = cell_of(VV(v), h);
            vcell3[3 * v] = cc.x; vcell3[3 * v + 1] = cc.y; vcell3[3 * v + 2] = cc.z;
        }
        wp.mark(5);
        const double tol2 = h * h;
        int nh = 0;
        const int ntest = ne * nv;
        for (int kb = 0; kb < ntest; kb += W) {
            const int k = kb + ln;
            bool hit = false;
            if (k < ntest) {
                const int e = k / nv, v = k - (k / nv) * nv;
                const int a = E[2 * e], b = E[2 * e + 1];
                if (v != a && v != b) {
                    const V3 pa = VV(a), pb = VV(b);
                    V3 mn = v3(dmin(pa.x, pb.x) - h, dmin(pa.y, pb.y) - h, dmin(pa.z, pb.z) - h);
                    V3 mx = v3(dmax(pa.x, pb.x) + h, dmax(pa.y, pb.y) + h, dmax(pa.z, pb.z) + h);
                    const I3 c0 = cell_of(mn, h), c1 = cell_of(mx, h);
                    int cx = vcell3[3 * v], cy = vcell3[3 * v + 1], cz = vcell3[3 * v + 2];
                    if (!(cx < c0.x || cx > c1.x || cy < c0.y || cy > c1.y || cz < c0.z || cz > c1.z)) {
                        double d2, t;
                        pt_seg(VV(v), pa, pb, d2, t);
                        hit = d2 <= tol2 && t > 1e-6 && t < 1.0 - 1e-6;
                    }
                }
            }
            uint64_t m = wp.ballot(hit);
            while (m) {
                const int kk = kb + ctz64(m);
                m &= m - 1;
                const int e = kk / nv, v = kk - (kk / nv) * nv;
                const int a = E[2 * e], b = E[2 * e + 1];
                const V3 pa = VV(a), pb = VV(b);
                V3 mn = v3(dmin(pa.x, pb.x) - h, dmin(pa.y, pb.y) - h, dmin(pa.z, pb.z) - h);
                V3 mx = v3(dmax(pa.x, pb.x) + h, dmax(pa.y, pb.y) + h, dmax(pa.z, pb.z) + h);
                const I3 c0 = cell_of(mn, h), c1 = cell_of(mx, h);
                double d2, t;
                pt_seg(VV(v), pa, pb, d2, t);
                if (nh >= HC) { out.status |= TOPO_E2_OVERFLOW; return out; }
                // the reference scans cells ix, iy, iz over [c0, c1]; key = scan index
                int cx = vcell3[3 * v], cy = vcell3[3 * v + 1], cz = vcell3[3 * v + 2];
                int nyz = (c1.y - c0.y + 1) * (c1.z - c0.z + 1);
                spt[nh] = t; spv[nh] = v; spe[nh] = e;
                spk[nh] = (cx - c0.x) * nyz + (cy - c0.y) * (c1.z - c0.z + 1) + (cz - c0.z);
                ++nh;
            }
        }
        wp.mark(6);
        int ne2 = 0;
        int h0 = 0;
        for (int e = 0; e < ne; ++e) {
            const int a = E[2 * e], b = E[2 * e + 1];
            int h1 = h0;
            while (h1 < nh && spe[h1] == e) ++h1;
            const int nsp = h1 - h0;
            if (nsp == 0) {
                if (ne2 + 1 > E2c) { out.status |= TOPO_E2_OVERFLOW; return out; }
                E2[2 * ne2] = a; E2[2 * ne2 + 1] = b; ++ne2;
                continue;
            }
            SplitArr sa; sa.t = spt.at(h0); sa.v = spv.at(h0); sa.k = spk.at(h0);
            // bring candidates into the reference's scan order, then its std::sort by t
            SplitByScan byscan;
            for (int i2 = 1; i2 < nsp; ++i2) {
                Split x = sa.get(i2);
                int j2 = i2 - 1;
                while (j2 >= 0 && byscan(x, sa.get(j2))) { sa.set(j2 + 1, sa.get(j2)); --j2; }
                sa.set(j2 + 1, x);
            }
            SplitByT byt;
            ref_sort(sa, nsp, byt);
            int prevv = a, last = -1;
            for (int k = 0; k < nsp; ++k) {
                int v = spv[h0 + k];
                if (v == last) continue;
                last = v;
                if (prevv != v) {
                    if (ne2 + 1 > E2c) { out.status |= TOPO_E2_OVERFLOW; return out; }
                    E2[2 * ne2] = prevv; E2[2 * ne2 + 1] = v; ++ne2;
                }
                prevv = v;
            }
            if (prevv != b) {
                if (ne2 + 1 > E2c) { out.status |= TOPO_E2_OVERFLOW; return out; }
                E2[2 * ne2] = prevv; E2[2 * ne2 + 1] = b; ++ne2;
            }
            h0 = h1;
        }
        wp.mark(7);
        // extract_loops
        if (nv > S2) { out.status |= TOPO_E2_OVERFLOW; return out; }
        for (int v = 0; v < nv; ++v) deg[v] = 0;
        for (int e = 0; e < ne2; ++e) { deg[E2[2 * e]] += 1; deg[E2[2 * e + 1]] += 1; }
        int acc = 0;
        for (int v = 0; v < nv; ++v) { aoff[v] = acc; afill[v] = acc; acc += deg[v]; }
        for (int e = 0; e < ne2; ++e) {
            int a = E2[2 * e], b = E2[2 * e + 1];
            int pa = afill[a]++; adjn[pa] = b; adje[pa] = e;
            int pb = afill[b]++; adjn[pb] = a; adje[pb] = e;
        }
        for (int e = 0; e < ne2; ++e) used[e] = 0;
        for (int v = 0; v < nv; ++v) sord[v] = v;
        StartOrder so; so.deg = deg;
        ref_sort(sord, nv, so);
        int nloops_here = 0;
        for (int si = 0; si < nv; ++si) {
            int sv = sord[si];
            for (int k = aoff[sv]; k < aoff[sv] + deg[sv]; ++k) {
                int ei = adje[k];
                if (used[ei]) continue;
                int plen = 0;
                path[plen++] = sv;
                int prev = sv, cur = adjn[k];
                used[ei] = 1;
                path[plen++] = cur;
                int steps = 0;
                bool closed = false;
                while (steps < 100000) {
                    steps += 1;
                    if (cur == sv) { closed = plen > 1; break; }
                    int pick = -1, first = -1;
                    for (int m = aoff[cur]; m < aoff[cur] + deg[cur]; ++m) {
                        if (used[adje[m]]) continue;
                        if (first < 0) first = m;
                        if (adjn[m] != prev) { pick = m; break; }
                    }
                    if (first < 0) break;
                    if (pick < 0) pick = first;
                    used[adje[pick]] = 1;
                    prev = cur;
                    cur = adjn[pick];
                    if (plen > E2c) { out.status |= TOPO_E2_OVERFLOW; return out; }
                    path[plen++] = cur;
                }
                if (!closed) continue;
                --plen;  // drop the repeated start vertex
                const int raw = out.loops_raw + nloops_here;
                ++nloops_here;
                if (plen < 3) continue;
                // accumulate_Sn_Gn_from_polyline
                V3 Sn = v3(0.0, 0.0, 0.0), Gn = v3(0.0, 0.0, 0.0);
                for (int m = 0; m < plen; ++m) {
                    V3 p0 = VV(path[m]);
                    V3 p1 = VV(path[(m + 1) % plen]);
                    Sn = add(Sn, mul(cross(p0, p1), 0.5));
                    V3 dx = sub(p1, p0);
                    double s = dot(p0, p1) + (1.0 / 3.0) * dot(dx, dx);
                    Gn = add(Gn, mul(dx, -(1.0 / 3.0) * s));
                }
                // contact_point_xc0
                double area = norm(Sn);
                if (area < 1e-14) continue;
                V3 nA = mul(Sn, -1.0 / area);
                V3 xc = dvd(cross(nA, Gn), area);
                if (area < area_eps) continue;
                if (dot(nA, dirAB) < 0.0) nA = mul(nA, -1.0);
                int st = sink(out.loops, ci, raw, Sn, Gn, nA, xc, area, plen, path, Vs);
                out.status |= st;
                out.loops += 1;
            }
        }
        out.loops_raw += nloops_here;
        wp.sync();
        wp.mark(8);
    }
    return out;
}

// ------------------------------------------------------------ walls
struct V2 { double x, y; };
ZDX_HD inline V2 v2(double x, double y) { V2 r; r.x = x; r.y = y; return r; }
ZDX_HD inline double cr2(const V2& a, const V2& b) { return a.x * b.y - a.y * b.x; }
ZDX_HD inline V2 sub2(const V2& a, const V2& b) { return v2(a.x - b.x, a.y - b.y); }
// is_inside_clip / compute_intersection_clip main.cpp:1072-1090
ZDX_HD inline bool inside_clip(const V2& p, const V2& c1, const V2& c2) { return cr2(sub2(c2, c1), sub2(p, c1)) >= 0.0; }
ZDX_HD inline V2 isect_clip(const V2& s, const V2& e, const V2& c1, const V2& c2) {
    V2 dc = sub2(c1, c2), dp = sub2(s, e);
    double n1 = cr2(c1, c2), n2 = cr2(s, e), n3 = cr2(dc, dp);
    if (fabs(n3) < 1e-30) return s;
    return v2((n1 * dp.x - n2 * dc.x) / n3, (n1 * dp.y - n2 * dc.y) / n3);
}
// build_basis main.cpp:1149-1157
ZDX_HD inline void basis(const V3& n, V3& u, V3& v) {
    if (fabs(n.x) < 0.9) u = normalize(cross(n, v3(1.0, 0.0, 0.0)));
    else u = normalize(cross(n, v3(0.0, 1.0, 0.0)));
    v = cross(n, u);
}

// compute_wall_contact main.cpp:1205-1258 for one wall triangle, run by the
// W lanes of a lane policy: the particle's world vertices are already
// computed (the reference re-transforms the mesh per wall triangle;
// quat_rotate is a pure function of the vertex, so the cached world vertices
// are the identical doubles).  The plane section (get_particle_plane_section
// 1172-1201) tests the (triangle, edge) slots lane-parallel; crossings are
// taken in slot order from the ballot mask and kept unless an earlier kept
// point lies within tol (unique_points 561-577), the kept-list check again
// lane-parallel.  Clip, area and centroid run redundantly on every lane.
// wv: world vertex array (3 doubles each); tri: vertex indices (3 per tri).
// Scratch (shared by the lanes): pts (3 cap_pts doubles), poly buffers A/B
// (2 (cap_pts + 8) doubles each).
template <class WP>
ZDX_HD inline bool wall_contact(const WP& wp, const double* wv, const int* tri, int T, const V3* wt, const V3& wn,
                                double tol, SArr<double> pts, SArr<double> PA, SArr<double> PB, int cap_pts,
                                double& area_out, V3& xc_out, V3& n_out, int& status) {
    const int W = WP::W;
    const int ln = wp.lane();
    V3 pn = wn;
    double pd = -dot(pn, wt[0]);
    const double t2 = tol * tol;
    int np = 0;
    const int nslots = 3 * T;
    for (int base = 0; base < nslots; base += W) {
        const int f = base + ln;
        bool hit = false;
        V3 ip = v3(0.0, 0.0, 0.0);
        if (f < nslots) {
            const int t = f / 3, i = f - 3 * t;
            const int i0 = tri[3 * t + i], i1 = tri[3 * t + (i + 1) % 3];
            V3 p0 = v3(wv[3 * i0], wv[3 * i0 + 1], wv[3 * i0 + 2]);
            V3 p1 = v3(wv[3 * i1], wv[3 * i1 + 1], wv[3 * i1 + 2]);
            double d0 = dot(pn, p0) + pd, d1 = dot(pn, p1) + pd;
            if (d0 * d1 < 0.0) {
                double tt = d0 / (d0 - d1);
                tt = dmin(1.0, dmax(0.0, tt));
                ip = add(p0, mul(sub(p1, p0), tt));
                hit = true;
            }
        }
        uint64_t m = wp.ballot(hit);
        while (m) {
            const int src = ctz64(m);
            m &= m - 1;
            const V3 q = v3(wp.shfl(ip.x, src), wp.shfl(ip.y, src), wp.shfl(ip.z, src));
            bool dup = false;
            for (int kb = 0; kb < np && !dup; kb += W) {
                const int k = kb + ln;
                bool near = false;
                if (k < np) near = norm2(sub(q, v3(pts[3 * k], pts[3 * k + 1], pts[3 * k + 2]))) <= t2;
                dup = wp.ballot(near) != 0ull;
            }
            if (!dup) {
                if (np >= cap_pts) { status = 1; return false; }
                pts[3 * np] = q.x; pts[3 * np + 1] = q.y; pts[3 * np + 2] = q.z;
                ++np;
            }
        }
    }
    if (np < 3) return false;
    V3 bu, bv;
    basis(pn, bu, bv);
    V3 o = wt[0];
    for (int m = 0; m < np; ++m) {
        V3 d = sub(v3(pts[3 * m], pts[3 * m + 1], pts[3 * m + 2]), o);
        PA[2 * m] = dot(d, bu);
        PA[2 * m + 1] = dot(d, bv);
    }
    V2 clip[3];
    for (int k = 0; k < 3; ++k) { V3 d = sub(wt[k], o); clip[k] = v2(dot(d, bu), dot(d, bv)); }
    // sutherland_hodgman_clip 1093-1122
    int nin = np;
    SArr<double> in = PA, outp = PB;
    V2 c1 = clip[2];
    for (int ck = 0; ck < 3; ++ck) {
        V2 c2 = clip[ck];
        if (nin == 0) break;
        int nout = 0;
        V2 sv = v2(in[2 * (nin - 1)], in[2 * (nin - 1) + 1]);
        for (int m = 0; m < nin; ++m) {
            V2 e = v2(in[2 * m], in[2 * m + 1]);
            if (inside_clip(e, c1, c2)) {
                if (!inside_clip(sv, c1, c2)) {
                    V2 x = isect_clip(sv, e, c1, c2);
                    if (nout >= cap_pts + 6) { status = 1; return false; }
                    outp[2 * nout] = x.x; outp[2 * nout + 1] = x.y; ++nout;
                }
                if (nout >= cap_pts + 6) { status = 1; return false; }
                outp[2 * nout] = e.x; outp[2 * nout + 1] = e.y; ++nout;
            } else if (inside_clip(sv, c1, c2)) {
                V2 x = isect_clip(sv, e, c1, c2);
                if (nout >= cap_pts + 6) { status = 1; return false; }
                outp[2 * nout] = x.x; outp[2 * nout + 1] = x.y; ++nout;
            }
            sv = e;
        }
        SArr<double> tmp = in; in = outp; outp = tmp;
        nin = nout;
        c1 = c2;
    }
    if (nin < 3) return false;
    // polygon_area_2d / polygon_centroid_2d 1125-1146
    double a = 0.0;
    for (int m = 0; m < nin; ++m) {
        int j = (m + 1) % nin;
        a += cr2(v2(in[2 * m], in[2 * m + 1]), v2(in[2 * j], in[2 * j + 1]));
    }
    a = fabs(a) * 0.5;
    if (a < tol * tol * 0.01) return false;
    double cx = 0.0, cy = 0.0;
    for (int m = 0; m < nin; ++m) { cx += in[2 * m]; cy += in[2 * m + 1]; }
    double nn = (double)nin;
    V2 c = v2(cx / nn, cy / nn);
    area_out = a;
    xc_out = add(add(o, mul(bu, c.x)), mul(bv, c.y));
    n_out = pn;
    return true;
}

// ------------------------------------------------------------ force laws
struct Mat {
    double eqr, young, poisson, mu, rest, inv_mass, mass;
};
struct Contact {
    V3 f;     // force on B (pair) / on the particle (wall)
    V3 r1;    // lever arm on A (pair) / on the particle (wall)
    V3 r2;    // lever arm on B (pair only)
};

// particle-particle force, main.cpp:2220-2278; ds is the pair's tangential
// spring, read-modify-written in loop order.
ZDX_HD inline Contact pp_force(const V3& nA, double area, const V3& xc, const Body& a, const Body& b, const Mat& ma,
                              const Mat& mb, double dt, V3& ds) {
    V3 rA = sub(xc, a.pos), rB = sub(xc, b.pos);
    V3 vA = add(a.vel, cross(a.omega, rA)), vB = add(b.vel, cross(b.omega, rB));
    V3 vrel = sub(vB, vA);
    double vn = dot(vrel, nA);
    double R1 = dmax(ma.eqr, 1e-12), R2 = dmax(mb.eqr, 1e-12);
    double E1 = dmax(ma.young, 1e-12), E2 = dmax(mb.young, 1e-12);
    double nu1 = ma.poisson, nu2 = mb.poisson;
    double kn = (E1 / R1 + E2 / R2) * area;
    double kt = (E1 / (2.0 * (1.0 + nu1) * R1) + E2 / (2.0 * (1.0 + nu2) * R2)) * area;
    double meff = 1.0 / (ma.inv_mass + mb.inv_mass);
    double e = dmin(ma.rest, mb.rest);
    e = dmin(0.9999, dmax(1e-6, e));
    double le = log(e);
    double ct = 2.0 * le * sqrt(kt * meff) / sqrt(le * le + 3.141592653589793 * 3.141592653589793);
    ct = fabs(ct);
    double dV = pow(area, 1.5);
    double fn = kn * cbrt(dV);
    if (fn < 0.0) fn = 0.0;
    V3 fnv = mul(nA, fn);
    V3 vt = sub(vrel, mul(nA, vn));
    ds = add(ds, mul(vt, dt));
    ds = sub(ds, mul(nA, dot(ds, nA)));
    V3 ft = add(mul(ds, -kt), mul(vt, -ct));
    double mu = dmin(ma.mu, mb.mu);
    double ftn = norm(ft), ftm = mu * fn;
    if (ftn > ftm && ftn > 1e-14) {
        ft = mul(ft, ftm / ftn);
        if (kt > 1e-12) ds = mul(ft, -1.0 / kt);
    }
    Contact c;
    c.f = add(fnv, ft);
    c.r1 = rA;
    c.r2 = rB;
    return c;
}

// particle-wall force, main.cpp:2363-2431 (normal damping on)
ZDX_HD inline Contact pw_force(const V3& n_in, double area, const V3& xc, const Body& p, const Mat& m, double wall_mu,
                              double dt, V3& ds) {
    V3 cn = n_in;
    V3 tp = sub(p.pos, xc);
    if (dot(cn, tp) < 0.0) cn = mul(cn, -1.0);
    V3 rP = sub(xc, p.pos);
    V3 vrel = add(p.vel, cross(p.omega, rP));
    double vn = dot(vrel, cn);
    double Rp = dmax(m.eqr, 1e-12), Ep = dmax(m.young, 1e-12), nu = m.poisson;
    double kn = (Ep / Rp) * area;
    double kt = (Ep / (2.0 * (1.0 + nu) * Rp)) * area;
    double meff = m.mass;
    double e = m.rest;
    e = dmin(0.9999, dmax(1e-6, e));
    double le = log(e);
    double cnn = 2.0 * le * sqrt(kn * meff) / sqrt(le * le + 3.141592653589793 * 3.141592653589793);
    double ct = 2.0 * le * sqrt(kt * meff) / sqrt(le * le + 3.141592653589793 * 3.141592653589793);
    cnn = fabs(cnn);
    ct = fabs(ct);
    double dV = pow(area, 1.5);
    double fe = kn * cbrt(dV);
    double fd = -cnn * vn;
    double fn = fe + fd;
    if (fn < 0.0) fn = 0.0;
    V3 fnv = mul(cn, fn);
    V3 vt = sub(vrel, mul(cn, vn));
    ds = add(ds, mul(vt, dt));
    ds = sub(ds, mul(cn, dot(ds, cn)));
    V3 ft = add(mul(ds, -kt), mul(vt, -ct));
    double mu = dmin(m.mu, wall_mu);
    double ftn = norm(ft), ftm = mu * fn;
    if (ftn > ftm && ftn > 1e-14) {
        ft = mul(ft, ftm / ftn);
        if (kt > 1e-12) ds = mul(ft, -1.0 / kt);
    }
    Contact c;
    c.f = add(fnv, ft);
    c.r1 = rP;
    c.r2 = v3(0.0, 0.0, 0.0);
    return c;
}

}  // namespace zdx
