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
// tri_tri_intersection_segment main.cpp:579-631, unique_points 561-577 inlined.
// The candidates are deduplicated as they come (a candidate is kept unless it
// lies within 1e-9 of an already kept one -- what unique_points does over the
// candidate list in order), and the kept points live in a fixed array that is
// only ever indexed by unrolled constants, so it stays in registers.  With
// codes != nullptr the provenance of the two endpoints is returned: r =
// candidate index (0..2: edge r of A crossing B's plane, 3..5: edge r-3 of B
// crossing A's plane), + 8 when the edge parameter was clamped to 0, + 16 when
// to 1.
ZDX_HD inline bool seg_plane_c(const V3& p0, const V3& p1, const V3& n, double d, V3& out, int& clamp) {
    V3 u = sub(p1, p0);
    double den = dot(n, u);
    double num = -(dot(n, p0) + d);
    if (fabs(den) < 1e-30) return false;
    double t = num / den;
    if (t < -1e-10 || t > 1.0 + 1e-10) return false;
    clamp = t < 0.0 ? 8 : (t > 1.0 ? 16 : 0);
    t = dmin(1.0, dmax(0.0, t));
    out = add(p0, mul(u, t));
    return true;
}
ZDX_HD inline bool tri_tri_impl(const V3* A, const V3& nA, double dA, const V3* B, const V3& nB, double dB, V3& o0,
                                V3& o1, int* codes) {
    V3 tau = cross(nA, nB);
    double lt = norm(tau);
    if (lt < 1e-20) return false;
    tau = dvd(tau, lt);
    V3 c[6];
    int cc[6];
    int nc = 0;
    const double t2 = 1e-9 * 1e-9;
    auto take = [&](const V3& p, int code) {
        bool keep = true;
#pragma unroll
        for (int j = 0; j < 6; ++j)
            if (j < nc && keep && norm2(sub(p, c[j])) <= t2) keep = false;
        if (!keep) return;
#pragma unroll
        for (int j = 0; j < 6; ++j)
            if (j == nc) { c[j] = p; cc[j] = code; }
        ++nc;
    };
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        V3 p;
        int cl = 0;
        if (seg_plane_c(A[i], A[(i + 1) % 3], nB, dB, p, cl) && point_in_tri(p, B, nB)) take(p, i | cl);
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        V3 p;
        int cl = 0;
        if (seg_plane_c(B[i], B[(i + 1) % 3], nA, dA, p, cl) && point_in_tri(p, A, nA)) take(p, (3 + i) | cl);
    }
    if (nc < 2) return false;
    double mn = INFINITY, mx = -INFINITY;
    int i0 = -1, i1 = -1, k0 = 0, k1 = 0;
    V3 p0 = v3(0.0, 0.0, 0.0), p1 = v3(0.0, 0.0, 0.0);
#pragma unroll
    for (int j = 0; j < 6; ++j) {
        if (j < nc) {
            const double t = dot(c[j], tau);
            if (t < mn) { mn = t; i0 = j; p0 = c[j]; k0 = cc[j]; }
            if (t > mx) { mx = t; i1 = j; p1 = c[j]; k1 = cc[j]; }
        }
    }
    if (i0 == i1) return false;
    if (norm2(sub(p1, p0)) < 1e-18) return false;
    if (dot(sub(p1, p0), tau) < 0.0) { V3 t = p0; p0 = p1; p1 = t; int k = k0; k0 = k1; k1 = k; }
    o0 = p0; o1 = p1;
    if (codes) { codes[0] = k0; codes[1] = k1; }
    return true;
}
ZDX_HD inline bool tri_tri(const V3* A, const V3& nA, double dA, const V3* B, const V3& nB, double dB, V3& o0,
                           V3& o1) {
    return tri_tri_impl(A, nA, dA, B, nB, dB, o0, o1, nullptr);
}
ZDX_HD inline bool tri_tri_prov(const V3* A, const V3& nA, double dA, const V3* B, const V3& nB, double dB,
                                V3& o0, V3& o1, int& code0, int& code1) {
    int k[2] = {0, 0};
    const bool r = tri_tri_impl(A, nA, dA, B, nB, dB, o0, o1, k);
    code0 = k[0];
    code1 = k[1];
    return r;
}

// Sensitivity of a segment endpoint to rigid motions of the two bodies.
// Motion parameters xi = (dA, wA, dB, wB): body X moves its material points
// by u_X(x) = d_X + w_X x (x - c_X).  An endpoint on edge (a, b) of body X
// crossing the plane (n) of a triangle of body Y moves by
//   dp = u_X(p) + e [n . (u_Y(p) - u_X(p))] / (n . e),   e = b - a
// (a clamped endpoint is the edge's vertex and moves with X).  J: 3 x 12,
// row-major.
ZDX_HD inline void endpoint_jacobian(int code, const V3* A, const V3* B, const V3& nA, const V3& nB, const V3& p,
                                     const V3& cA, const V3& cB, double* J) {
    const int r = code & 7;
    const bool onA = r < 3;                   // the edge belongs to A
    const int ei = onA ? r : r - 3;
    const V3* T = onA ? A : B;
    const V3 e = sub(T[(ei + 1) % 3], T[ei]);
    const V3 n = onA ? nB : nA;               // plane of the other body's triangle
    const int cx = onA ? 0 : 6, cy = onA ? 6 : 0;
    const V3 rx = sub(p, onA ? cA : cB), ry = sub(p, onA ? cB : cA);
    for (int k = 0; k < 36; ++k) J[k] = 0.0;
    // u_X(p) = d_X - [rx]x w_X
    auto put_rigid = [&](int col, const V3& rr, double s) {
        // rows: x, y, z; translation block I, rotation block -[rr]x
        J[0 * 12 + col + 0] += s; J[1 * 12 + col + 1] += s; J[2 * 12 + col + 2] += s;
        J[0 * 12 + col + 4] += s * rr.z;  J[0 * 12 + col + 5] -= s * rr.y;
        J[1 * 12 + col + 3] -= s * rr.z;  J[1 * 12 + col + 5] += s * rr.x;
        J[2 * 12 + col + 3] += s * rr.y;  J[2 * 12 + col + 4] -= s * rr.x;
    };
    put_rigid(cx, rx, 1.0);
    if (code & 24) return;
    const double ne = dot(n, e);
    // row vector g = n^T (J_Y - J_X) / (n . e); J += e g
    double g[12];
    for (int k = 0; k < 12; ++k) g[k] = 0.0;
    auto row_rigid = [&](int col, const V3& rr, double s) {
        // n^T [I, -[rr]x] = [n, (rr x n)^T]  (n . (w x rr) = w . (rr x n))
        const V3 rn = cross(rr, n);
        g[col + 0] += s * n.x; g[col + 1] += s * n.y; g[col + 2] += s * n.z;
        g[col + 3] += s * rn.x; g[col + 4] += s * rn.y; g[col + 5] += s * rn.z;
    };
    row_rigid(cy, ry, 1.0 / ne);
    row_rigid(cx, rx, -1.0 / ne);
    for (int k = 0; k < 12; ++k) {
        J[0 * 12 + k] += e.x * g[k];
        J[1 * 12 + k] += e.y * g[k];
        J[2 * 12 + k] += e.z * g[k];
    }
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
// The same order on (key, vertex) elements: key = INT_MIN for degree 2, else
// -degree, so key a < key b exactly when StartOrder puts a before b; the sort
// then moves the key with its vertex (the same permutation, one load per
// element instead of a vertex and then its degree).
struct KeyId { int k, v; };
struct KeyIdArr {
    using value_type = KeyId;
    int* k;
    int* v;
    ZDX_HD KeyId get(int i) const { KeyId r; r.k = k[i]; r.v = v[i]; return r; }
    ZDX_HD void set(int i, const KeyId& x) { k[i] = x.k; v[i] = x.v; }
};
struct ByKey { ZDX_HD bool operator()(const KeyId& a, const KeyId& b) const { return a.k < b.k; } };
ZDX_HD inline int start_key(int deg) { return deg == 2 ? (-2147483647 - 1) : -deg; }
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
    uint64_t shfl64(uint64_t v, int) const { return v; }
    int atomic_cas(int* p, int expected, int desired) const {
        const int old = *p;
        if (old == expected) *p = desired;
        return old;
    }
    void atomic_min(int* p, int v) const { if (v < *p) *p = v; }
    void atomic_add(int* p, int v) const { *p += v; }
    int readlane(int v, int) const { return v; }
    uint64_t readlane64(uint64_t v, int) const { return v; }
    int scan_excl(int v, int& total) const { total = v; return 0; }   // exclusive prefix over the lanes
    void sync() const {}
    void mark(int) const {}   // phase boundary (timed by profiling builds of the device policy)
    void count(int) const {}  // work counter of profiling builds
};

// index of the lowest set bit (m != 0)
ZDX_HD inline int ctz64(uint64_t m) { return __builtin_ctzll(m); }

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
// V: the number of world vertices (every triangle corner is one of them).
// srcs (optional, cap_pts entries): the (triangle, edge) slot 3 t + e each
// kept section point came from, for wall_contact_jac.
template <class WP>
ZDX_HD inline bool wall_contact(const WP& wp, const double* wv, const int* tri, int T, int V, const V3* wt,
                                const V3& wn, double tol, SArr<double> pts, SArr<double> PA, SArr<double> PB,
                                int cap_pts, double& area_out, V3& xc_out, V3& n_out, int& status,
                                int32_t* srcs = nullptr) {
    const int W = WP::W;
    const int ln = wp.lane();
    V3 pn = wn;
    double pd = -dot(pn, wt[0]);
    const double t2 = tol * tol;
    int np = 0;
    // An edge crosses the plane only when its endpoints' signed distances
    // (computed below exactly as here) have a negative product, so only when
    // some vertex lies strictly below and some strictly above: without both,
    // the section is empty and the contact is rejected (np < 3) unchanged.
    {
        bool below = false, above = false;
        for (int vb = 0; vb < V; vb += W) {
            const int v = vb + ln;
            if (v < V) {
                const double d = dot(pn, v3(wv[3 * v], wv[3 * v + 1], wv[3 * v + 2])) + pd;
                below = below || d < 0.0;
                above = above || d > 0.0;
            }
            if (wp.ballot(below) != 0ull && wp.ballot(above) != 0ull) break;
        }
        if (!(wp.ballot(below) != 0ull && wp.ballot(above) != 0ull)) return false;
    }
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
                if (srcs) srcs[np] = base + src;
                ++np;
            }
        }
    }
    if (srcs) srcs[cap_pts] = np;   // the count, after the slots
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

// Derivatives of a wall contact (area and contact point of wall_contact)
// with respect to a rigid motion xi = (d, w) of the particle: every world
// vertex p moves by d + w x (p - x).  One thread; it repeats wall_contact's
// arithmetic on the kept section points (srcs: their edge slots and count,
// as wall_contact recorded them) with six tangents carried beside each value,
// so every branch (clip tests, area sign) is decided on the same doubles.
// A kept point p = p0 + t (p1 - p0), t = d0 / (d0 - d1) on the edge (p0,
// p1) with signed distances d0, d1 to the wall plane, has the tangent
// dp0 + t (dp1 - dp0) + (p1 - p0) dt, dt = (d0 dd1 - d1 dd0) / (d0 - d1)^2.
// Scratch T: 14 (cap_pts + 8) doubles per polygon buffer, two buffers.
// dA: 6 doubles; dX: 3 x 6 (row-major).  Returns false when wall_contact
// would (no contact).
ZDX_HD inline bool wall_contact_jac(const double* wv, const int* tri, const V3& x, const V3* wt, const V3& wn,
                                    double tol, const int32_t* srcs, int cap_pts, SArr<double> TA, SArr<double> TB,
                                    double* dA, double* dX) {
    const int np = srcs[cap_pts];
    if (np < 3) return false;
    V3 pn = wn;
    double pd = -dot(pn, wt[0]);
    V3 bu, bv;
    basis(pn, bu, bv);
    V3 o = wt[0];
    // polygon vertex m: [0] u, [1] v, [2..8) du/dxi, [8..14) dv/dxi
    for (int m = 0; m < np; ++m) {
        const int f = srcs[m];
        const int t = f / 3, i = f - 3 * t;
        const int i0 = tri[3 * t + i], i1 = tri[3 * t + (i + 1) % 3];
        V3 p0 = v3(wv[3 * i0], wv[3 * i0 + 1], wv[3 * i0 + 2]);
        V3 p1 = v3(wv[3 * i1], wv[3 * i1 + 1], wv[3 * i1 + 2]);
        double d0 = dot(pn, p0) + pd, d1 = dot(pn, p1) + pd;
        double tt = d0 / (d0 - d1);
        const bool clamped = tt < 0.0 || tt > 1.0;
        tt = dmin(1.0, dmax(0.0, tt));
        V3 p = add(p0, mul(sub(p1, p0), tt));
        V3 e = sub(p1, p0);
        V3 r0 = sub(p0, x), r1 = sub(p1, x);
        V3 dd = sub(p, o);
        TA[14 * m] = dot(dd, bu);
        TA[14 * m + 1] = dot(dd, bv);
        const double den = (d0 - d1) * (d0 - d1);
        for (int j = 0; j < 6; ++j) {
            V3 a = j < 3 ? v3(j == 0, j == 1, j == 2) : cross(v3(j == 3, j == 4, j == 5), r0);
            V3 b = j < 3 ? a : cross(v3(j == 3, j == 4, j == 5), r1);
            const double dtt = clamped ? 0.0 : (d0 * dot(pn, b) - d1 * dot(pn, a)) / den;
            V3 dp = add(add(a, mul(sub(b, a), tt)), mul(e, dtt));
            TA[14 * m + 2 + j] = dot(dp, bu);
            TA[14 * m + 8 + j] = dot(dp, bv);
        }
    }
    V2 clip[3];
    for (int k = 0; k < 3; ++k) { V3 d = sub(wt[k], o); clip[k] = v2(dot(d, bu), dot(d, bv)); }
    // sutherland_hodgman_clip 1093-1122 with tangents
    int nin = np;
    SArr<double> in = TA, outp = TB;
    V2 c1 = clip[2];
    auto emit = [&](int& nout, const double* src14) {
        for (int q = 0; q < 14; ++q) outp[14 * nout + q] = src14[q];
        ++nout;
    };
    for (int ck = 0; ck < 3; ++ck) {
        V2 c2 = clip[ck];
        if (nin == 0) break;
        int nout = 0;
        int sidx = nin - 1;
        for (int m = 0; m < nin; ++m) {
            V2 e = v2(in[14 * m], in[14 * m + 1]);
            V2 sv = v2(in[14 * sidx], in[14 * sidx + 1]);
            const bool ein = inside_clip(e, c1, c2), sin_ = inside_clip(sv, c1, c2);
            if (ein != sin_) {
                // the crossing of the edge (sv, e) with the clip line
                double rec[14];
                V2 dc = sub2(c1, c2), dp = sub2(sv, e);
                double n1 = cr2(c1, c2), n2 = cr2(sv, e), n3 = cr2(dc, dp);
                if (fabs(n3) < 1e-30) {
                    for (int q = 0; q < 14; ++q) rec[q] = in[14 * sidx + q];
                } else {
                    const double X = (n1 * dp.x - n2 * dc.x) / n3, Y = (n1 * dp.y - n2 * dc.y) / n3;
                    rec[0] = X;
                    rec[1] = Y;
                    for (int j = 0; j < 6; ++j) {
                        V2 dsv = v2(in[14 * sidx + 2 + j], in[14 * sidx + 8 + j]);
                        V2 de = v2(in[14 * m + 2 + j], in[14 * m + 8 + j]);
                        V2 ddp = sub2(dsv, de);
                        const double dn2 = cr2(dsv, e) + cr2(sv, de), dn3 = cr2(dc, ddp);
                        rec[2 + j] = (n1 * ddp.x - dn2 * dc.x - X * dn3) / n3;
                        rec[8 + j] = (n1 * ddp.y - dn2 * dc.y - Y * dn3) / n3;
                    }
                }
                if (nout >= cap_pts + 6) return false;
                emit(nout, rec);
            }
            if (ein) {
                if (nout >= cap_pts + 6) return false;
                double rec[14];
                for (int q = 0; q < 14; ++q) rec[q] = in[14 * m + q];
                emit(nout, rec);
            }
            sidx = m;
        }
        SArr<double> tmp = in; in = outp; outp = tmp;
        nin = nout;
        c1 = c2;
    }
    if (nin < 3) return false;
    double a = 0.0;
    double da[6] = {0, 0, 0, 0, 0, 0};
    for (int m = 0; m < nin; ++m) {
        int j = (m + 1) % nin;
        V2 pm = v2(in[14 * m], in[14 * m + 1]), pj = v2(in[14 * j], in[14 * j + 1]);
        a += cr2(pm, pj);
        for (int q = 0; q < 6; ++q)
            da[q] += cr2(v2(in[14 * m + 2 + q], in[14 * m + 8 + q]), pj) +
                     cr2(pm, v2(in[14 * j + 2 + q], in[14 * j + 8 + q]));
    }
    const double sg = a < 0.0 ? -0.5 : 0.5;
    a = fabs(a) * 0.5;
    if (a < tol * tol * 0.01) return false;
    const double nn = (double)nin;
    for (int q = 0; q < 6; ++q) {
        double su = 0.0, sv2 = 0.0;
        for (int m = 0; m < nin; ++m) { su += in[14 * m + 2 + q]; sv2 += in[14 * m + 8 + q]; }
        dA[q] = sg * da[q];
        V3 dxq = add(mul(bu, su / nn), mul(bv, sv2 / nn));
        dX[q] = dxq.x; dX[6 + q] = dxq.y; dX[12 + q] = dxq.z;
    }
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
// spring, read-modify-written in loop order.  damped: add the normal damping
// the reference computes and leaves commented out (main.cpp:2244-2248,
// 2258-2260), -cn vn while the bodies approach (vn < 0).
ZDX_HD inline Contact pp_force(const V3& nA, double area, const V3& xc, const Body& a, const Body& b, const Mat& ma,
                              const Mat& mb, double dt, V3& ds, bool damped = false) {
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
    double cn = 2.0 * le * sqrt(kn * meff) / sqrt(le * le + 3.141592653589793 * 3.141592653589793);
    double ct = 2.0 * le * sqrt(kt * meff) / sqrt(le * le + 3.141592653589793 * 3.141592653589793);
    cn = fabs(cn);
    ct = fabs(ct);
    double dV = pow(area, 1.5);
    double fn = kn * cbrt(dV);
    if (damped && vn < 0.0) fn += -cn * vn;
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

// Minimal-energy particle-particle force of the time step (force model
// "minimal"): polyline.py:655-701, Fn = k/2 |Sn| along the A->B normal nA,
// with no damping and no tangential force.  The lever arm is the loop
// centroid: xc0 = (nA x Gn)/|Sn| is -2/3 of the centroid's part normal to nA
// (Stokes), so -3/2 xc0 is the centroid up to a multiple of nA, which the
// torque of a force along nA does not see.
ZDX_HD inline Contact pp_force_minimal(const V3& nA, double area, const V3& xc0, const Body& a, const Body& b,
                                      double k) {
    V3 xc = mul(xc0, -1.5);
    Contact r;
    r.f = mul(nA, 0.5 * k * area);
    r.r1 = sub(xc, a.pos);
    r.r2 = sub(xc, b.pos);
    return r;
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
