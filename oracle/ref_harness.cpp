// ref_harness.cpp — exposes the compiled ZDEM-X reference (src/main.cpp) to the
// oracle checks.
//
// TEST INFRASTRUCTURE ONLY.  This file contains no reference code: it includes
// the reference source where it lies (path given by -DZDX_REF_MAIN=... from
// oracle/Makefile) with `main` renamed, so the reference's own static
// functions can be called, and it switches the state-dump precision of
// write_particle_state_txt (main.cpp:1851-1852) from 9 fixed decimals to 17
// significant digits in scientific form so dumped trajectories are exact.
// Output: oracle/_ref/libzdx_ref.so (never committed; built only where the
// reference is present).
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cctype>
#include <fstream>
#include <filesystem>
#include <iomanip>
#include <iostream>
#include <limits>
#include <random>
#include <chrono>
#include <sstream>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#define main zdem_reference_main
#define setprecision(n) setprecision(17)
#define fixed scientific
#include ZDX_REF_MAIN
#undef fixed
#undef setprecision
#undef main

namespace {
std::vector<std::array<Vec3, 3>> tris_from(const double* t, int T) {
    std::vector<std::array<Vec3, 3>> out(T);
    for (int i = 0; i < T; ++i)
        for (int k = 0; k < 3; ++k) out[i][k] = Vec3{t[9 * i + 3 * k], t[9 * i + 3 * k + 1], t[9 * i + 3 * k + 2]};
    return out;
}
void put(double* o, const Vec3& v) { o[0] = v.x; o[1] = v.y; o[2] = v.z; }
}  // namespace

extern "C" {

// zdem_cpu entry point (main.cpp:1872) with argv as given.
int ref_main(int argc, char** argv) { return zdem_reference_main(argc, argv); }

// load_stl (main.cpp:308)
int ref_load_stl(const char* path, double* tris, int cap) {
    std::vector<std::array<Vec3, 3>> t;
    if (!load_stl(path, t)) return -1;
    int T = static_cast<int>(t.size());
    if (tris && T <= cap)
        for (int i = 0; i < T; ++i)
            for (int k = 0; k < 3; ++k) put(tris + 9 * i + 3 * k, t[i][k]);
    return T;
}

// build_mesh (main.cpp:445): props as in orc_build_mesh
int ref_build_mesh(const double* tris, int T, int center, double* out_tris, double* props) {
    Mesh m = build_mesh(tris_from(tris, T), center != 0);
    if (out_tris)
        for (int i = 0; i < T; ++i)
            for (int k = 0; k < 3; ++k) put(out_tris + 9 * i + 3 * k, m.tris[i][k]);
    if (props) {
        props[0] = m.volume; put(props + 1, m.center);
        props[4] = m.radius; props[5] = m.mean_edge; props[6] = m.bbox_diag;
        for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) props[7 + 3 * i + j] = m.inertia_unit.m[i][j];
    }
    return static_cast<int>(m.vertices.size());
}

// transform_tris (main.cpp:499)
void ref_transform(const double* tris, int T, const double* pos, const double* quat, double* out) {
    Mesh m;
    m.tris = tris_from(tris, T);
    Transform tf;
    tf.pos = Vec3{pos[0], pos[1], pos[2]};
    tf.rot = Quat{quat[0], quat[1], quat[2], quat[3]};
    auto w = transform_tris(m, tf);
    for (int i = 0; i < T; ++i)
        for (int k = 0; k < 3; ++k) put(out + 9 * i + 3 * k, w[i][k]);
}

// plane_from_tri + tri_tri_intersection_segment (main.cpp:518, 579)
int ref_tri_tri(const double* a, const double* b, double* out) {
    auto A = tris_from(a, 1), B = tris_from(b, 1);
    Vec3 nA, nB;
    double dA = 0, dB = 0;
    if (!plane_from_tri(A[0], nA, dA)) return 0;
    if (!plane_from_tri(B[0], nB, dB)) return 0;
    Vec3 p0, p1;
    if (!tri_tri_intersection_segment(A[0], nA, dA, B[0], nB, dB, p0, p1)) return 0;
    put(out, p0); put(out + 3, p1);
    return 1;
}

// Narrow phase of one pair with the reference's own primitives: transform_tris
// (499), plane_from_tri (518), per-tri AABBs and the strict-compare reject
// (2093-2145), tri_tri_intersection_segment (579).  Returns the segment count
// (segments written up to cap, as p0,p1) and the tri_checks counter.
int ref_pair_segments(const double* trisA, int TA, const double* trisB, int TB,
                      const double* posA, const double* quatA, const double* posB, const double* quatB,
                      double* seg, int cap, int64_t* tri_checks) {
    Mesh mA, mB;
    mA.tris = tris_from(trisA, TA);
    mB.tris = tris_from(trisB, TB);
    Transform fa, fb;
    fa.pos = Vec3{posA[0], posA[1], posA[2]}; fa.rot = Quat{quatA[0], quatA[1], quatA[2], quatA[3]};
    fb.pos = Vec3{posB[0], posB[1], posB[2]}; fb.rot = Quat{quatB[0], quatB[1], quatB[2], quatB[3]};
    auto A = transform_tris(mA, fa), B = transform_tris(mB, fb);
    std::vector<Vec3> nA(TA), nB(TB);
    std::vector<double> dA(TA), dB(TB);
    for (int i = 0; i < TA; ++i) if (!plane_from_tri(A[i], nA[i], dA[i])) { nA[i] = Vec3{}; dA[i] = 0.0; }
    for (int i = 0; i < TB; ++i) if (!plane_from_tri(B[i], nB[i], dB[i])) { nB[i] = Vec3{}; dB[i] = 0.0; }
    auto lo = [](const std::array<Vec3, 3>& t) {
        return Vec3{std::min({t[0].x, t[1].x, t[2].x}), std::min({t[0].y, t[1].y, t[2].y}), std::min({t[0].z, t[1].z, t[2].z})};
    };
    auto hi = [](const std::array<Vec3, 3>& t) {
        return Vec3{std::max({t[0].x, t[1].x, t[2].x}), std::max({t[0].y, t[1].y, t[2].y}), std::max({t[0].z, t[1].z, t[2].z})};
    };
    int n = 0;
    int64_t checks = 0;
    for (int i = 0; i < TA; ++i) {
        if (norm2(nA[i]) < 1e-30) continue;
        Vec3 a0 = lo(A[i]), a1 = hi(A[i]);
        for (int j = 0; j < TB; ++j) {
            checks++;
            Vec3 b0 = lo(B[j]), b1 = hi(B[j]);
            if (a1.x < b0.x || a0.x > b1.x || a1.y < b0.y || a0.y > b1.y || a1.z < b0.z || a0.z > b1.z) continue;
            if (norm2(nB[j]) < 1e-30) continue;
            Vec3 p0, p1;
            if (tri_tri_intersection_segment(A[i], nA[i], dA[i], B[j], nB[j], dB[j], p0, p1)) {
                if (n < cap) { put(seg + 6 * n, p0); put(seg + 6 * n + 3, p1); }
                ++n;
            }
        }
    }
    *tri_checks = checks;
    return n;
}

// label_segments_by_contact_greedy (main.cpp:804): comp_of[s] = component index
// in sorted order, order_out = concatenated element order.
int ref_components(const double* seg, int ns, double tol, int32_t* comp_sizes, int32_t* order_out) {
    std::vector<std::pair<Vec3, Vec3>> s(ns);
    for (int i = 0; i < ns; ++i)
        s[i] = {Vec3{seg[6 * i], seg[6 * i + 1], seg[6 * i + 2]}, Vec3{seg[6 * i + 3], seg[6 * i + 4], seg[6 * i + 5]}};
    std::vector<std::vector<int>> comps;
    label_segments_by_contact_greedy(s, tol, comps);
    int k = 0;
    for (std::size_t c = 0; c < comps.size(); ++c) {
        comp_sizes[c] = static_cast<int32_t>(comps[c].size());
        for (int v : comps[c]) order_out[k++] = v;
    }
    return static_cast<int>(comps.size());
}

// snap_endpoints + split_t_junctions + extract_loops (main.cpp:680, 900, 963)
// on one component.  Writes loop vertex lists (vids) and snapped vertices.
int ref_loops(const double* seg, int ns, double tol, double* V_out, int* nv_out,
              int32_t* loop_len, int32_t* loop_vids, int cap) {
    std::vector<std::pair<Vec3, Vec3>> s(ns);
    for (int i = 0; i < ns; ++i)
        s[i] = {Vec3{seg[6 * i], seg[6 * i + 1], seg[6 * i + 2]}, Vec3{seg[6 * i + 3], seg[6 * i + 4], seg[6 * i + 5]}};
    std::vector<Vec3> V, V2;
    std::vector<std::pair<int, int>> E, E2;
    snap_endpoints(s, tol, V, E);
    split_t_junctions(V, E, tol, V2, E2);
    auto L = extract_loops(V2, E2);
    *nv_out = static_cast<int>(V2.size());
    for (std::size_t i = 0; i < V2.size() && static_cast<int>(i) < cap; ++i) put(V_out + 3 * i, V2[i]);
    int k = 0;
    for (std::size_t l = 0; l < L.size(); ++l) {
        loop_len[l] = static_cast<int32_t>(L[l].size());
        for (int v : L[l]) if (k < cap) loop_vids[k++] = v;
    }
    return static_cast<int>(L.size());
}

// accumulate_Sn_Gn_from_polyline + contact_point_xc0 (main.cpp:1027, 1042)
int ref_sn_xc0(const double* pts, int m, double* Sn, double* Gn, double* xc0, double* nA, double* area) {
    std::vector<Vec3> L(m);
    for (int i = 0; i < m; ++i) L[i] = Vec3{pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]};
    auto r = accumulate_Sn_Gn_from_polyline(L);
    put(Sn, r.first); put(Gn, r.second);
    Vec3 x, n;
    int ok = contact_point_xc0(r.first, r.second, x, n, *area) ? 1 : 0;
    put(xc0, x); put(nA, n);
    return ok;
}

// sutherland_hodgman_clip (main.cpp:1093)
int ref_sh_clip(const double* subj, int ns, const double* clip, int nc, double* out, int cap) {
    std::vector<Vec2> s(ns), c(nc);
    for (int i = 0; i < ns; ++i) s[i] = Vec2{subj[2 * i], subj[2 * i + 1]};
    for (int i = 0; i < nc; ++i) c[i] = Vec2{clip[2 * i], clip[2 * i + 1]};
    auto o = sutherland_hodgman_clip(s, c);
    int n = static_cast<int>(o.size());
    for (int i = 0; i < n && i < cap; ++i) { out[2 * i] = o[i].x; out[2 * i + 1] = o[i].y; }
    return n;
}

// compute_wall_contact (main.cpp:1205)
int ref_wall_contact(const double* tris_body, int T, const double* pos, const double* quat,
                     const double* wall_tri, const double* wall_n, double tol,
                     double* area, double* xc, double* n) {
    Mesh m;
    m.tris = tris_from(tris_body, T);
    Transform tf;
    tf.pos = Vec3{pos[0], pos[1], pos[2]};
    tf.rot = Quat{quat[0], quat[1], quat[2], quat[3]};
    auto w = tris_from(wall_tri, 1);
    Vec3 x, nn;
    if (!compute_wall_contact(m, tf, w[0], Vec3{wall_n[0], wall_n[1], wall_n[2]}, 0.0, tol, *area, x, nn)) return 0;
    put(xc, x); put(n, nn);
    return 1;
}

// integrate_particle (main.cpp:1494)
void ref_integrate(double* st, const double* Ibinv, double inv_mass, const double* F, const double* T,
                   const double* g, double dt) {
    Particle p;
    p.tf.pos = Vec3{st[0], st[1], st[2]};
    p.tf.rot = Quat{st[3], st[4], st[5], st[6]};
    p.vel = Vec3{st[7], st[8], st[9]};
    p.omega = Vec3{st[10], st[11], st[12]};
    p.L = Vec3{st[13], st[14], st[15]};
    p.inv_mass = inv_mass;
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) p.inertia_body_inv.m[i][j] = Ibinv[3 * i + j];
    integrate_particle(p, Vec3{F[0], F[1], F[2]}, Vec3{T[0], T[1], T[2]}, Vec3{g[0], g[1], g[2]}, dt);
    put(st, p.tf.pos);
    st[3] = p.tf.rot.w; st[4] = p.tf.rot.x; st[5] = p.tf.rot.y; st[6] = p.tf.rot.z;
    put(st + 7, p.vel); put(st + 10, p.omega); put(st + 13, p.L);
}

}  // extern "C"
