// zdx-pair: the per-pair tool of polyline.py (polyline.py:707-834) on the
// device contact pipeline.
//
//   zdx-pair stlA stlB [--tA x y z] [--qA w x y z] [--tB x y z] [--qB w x y z]
//            [--split-contacts] [--tol T] [--k K] [--vtk routeB_intersection.vtk]
//            [--outA A_transformed.vtk] [--outB B_transformed.vtk] [--device N]
//
// Writes the two transformed meshes and the contact loops (one closed polyline
// per loop, its contact point xc0 as a vertex, CELL_DATA contact_id) in
// polyline.py's formats, and prints its report.  The meshes are taken as given
// (no centring, as polyline.py) and moved by polyline.py's rotation formula;
// the loops come from zdx_contact_pairs on those moved triangles, i.e.
// the reference C++ pipeline (src/main.cpp:2061-2219) with --tol as the snap /
// adjacency tolerance, and the force is polyline.py's minimal one,
// Fn = k/2 |Sn|, F_A = Fn nA with nA = -Sn/|Sn| (polyline.py:655-701).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/zdx.h"
#include "zdx_geom.h"
#include "zdx_internal.h"

using zdx::py_float_repr;
using zdx::V3;

namespace {

struct Args {
    std::string stlA, stlB;
    double tA[3] = {0, 0, 0}, qA[4] = {1, 0, 0, 0}, tB[3] = {0, 0, 0}, qB[4] = {1, 0, 0, 0};
    bool split = false;
    double tol = 1e-6, k = 1.0;
    std::string vtk = "routeB_intersection.vtk", outA = "A_transformed.vtk", outB = "B_transformed.vtk";
    int device = 0;
};

void usage() {
    std::fprintf(stderr,
                 "usage: zdx-pair stlA stlB [--tA x y z] [--qA w x y z] [--tB x y z] [--qB w x y z] "
                 "[--split-contacts] [--tol T] [--k K] [--vtk F] [--outA F] [--outB F] [--device N]\n");
}

bool parse(int argc, char** argv, Args& a) {
    std::vector<std::string> pos;
    for (int i = 1; i < argc; ++i) {
        const std::string s = argv[i];
        auto nums = [&](double* out, int n) {
            if (i + n >= argc) return false;
            for (int k = 0; k < n; ++k) out[k] = std::strtod(argv[++i], nullptr);
            return true;
        };
        auto str = [&](std::string& out) {
            if (i + 1 >= argc) return false;
            out = argv[++i];
            return true;
        };
        bool ok = true;
        if (s == "--tA") ok = nums(a.tA, 3);
        else if (s == "--qA") ok = nums(a.qA, 4);
        else if (s == "--tB") ok = nums(a.tB, 3);
        else if (s == "--qB") ok = nums(a.qB, 4);
        else if (s == "--split-contacts") a.split = true;
        else if (s == "--tol") ok = nums(&a.tol, 1);
        else if (s == "--k") ok = nums(&a.k, 1);
        else if (s == "--vtk") ok = str(a.vtk);
        else if (s == "--outA") ok = str(a.outA);
        else if (s == "--outB") ok = str(a.outB);
        else if (s == "--device") { double d = 0; ok = nums(&d, 1); a.device = (int)d; }
        else if (s.size() > 2 && s[0] == '-' && s[1] == '-') ok = false;
        else pos.push_back(s);
        if (!ok) return false;
    }
    if (pos.size() != 2) return false;
    a.stlA = pos[0];
    a.stlB = pos[1];
    return true;
}

std::string vec(const double* v) {
    return "[" + py_float_repr(v[0]) + ", " + py_float_repr(v[1]) + ", " + py_float_repr(v[2]) + "]";
}

bool load(const std::string& path, std::vector<double>& tris) {
    int32_t T = 0;
    if (zdx_load_stl(path.c_str(), nullptr, 0, &T) != ZDX_OK && T <= 0) return false;
    tris.assign(9 * (size_t)T, 0.0);
    return zdx_load_stl(path.c_str(), tris.data(), T, &T) == ZDX_OK;
}

}  // namespace

int main(int argc, char** argv) {
    Args a;
    if (!parse(argc, argv, a)) {
        usage();
        return 2;
    }
    std::vector<double> trisA0, trisB0;
    if (!load(a.stlA, trisA0) || !load(a.stlB, trisB0)) {
        std::fprintf(stderr, "zdx-pair: failed to read %s or %s\n", a.stlA.c_str(), a.stlB.c_str());
        return 1;
    }
    const int64_t TA = (int64_t)trisA0.size() / 9, TB = (int64_t)trisB0.size() / 9;
    std::vector<double> trisA(trisA0.size()), trisB(trisB0.size());
    zdx::polyline_transform_points(trisA0.data(), 3 * TA, a.tA, a.qA, trisA.data());
    zdx::polyline_transform_points(trisB0.data(), 3 * TB, a.tB, a.qB, trisB.data());
    if (!zdx::write_vtk_polydata_tris(a.outA, trisA.data(), TA, "Mesh A (transformed)") ||
        !zdx::write_vtk_polydata_tris(a.outB, trisB.data(), TB, "Mesh B (transformed)")) {
        std::fprintf(stderr, "zdx-pair: cannot write %s / %s\n", a.outA.c_str(), a.outB.c_str());
        return 1;
    }
    std::printf("Transformed meshes written: %s, %s\n", a.outA.c_str(), a.outB.c_str());

    zdx_ctx* ctx = nullptr;
    if (zdx_create(&ctx, a.device) != ZDX_OK) {
        std::fprintf(stderr, "zdx-pair: no HIP device %d\n", a.device);
        return 1;
    }
    int32_t ma = 0, mb = 0;
    // the meshes go in already transformed, at the identity pose (an exact
    // map), so the loops are cut from the very triangles written above
    if (zdx_mesh_register(ctx, trisA.data(), (int32_t)TA, 0, &ma, nullptr) != ZDX_OK ||
        zdx_mesh_register(ctx, trisB.data(), (int32_t)TB, 0, &mb, nullptr) != ZDX_OK) {
        std::fprintf(stderr, "zdx-pair: %s\n", zdx_last_error(ctx));
        zdx_destroy(ctx);
        return 1;
    }
    zdx_pair_in in{};
    in.mesh_a = ma; in.mesh_b = mb;
    in.quat_a[0] = in.quat_b[0] = 1.0;
    zdx_contact_opts o{};
    o.split_contacts = a.split ? 1 : 0;
    o.mode = ZDX_MODE_BATCHED;
    o.k = a.k;
    o.tol = a.tol;
    std::vector<int32_t> lo(2), vo(1);
    std::vector<double> rec, vtx;
    int32_t ns = 0, nc = 0, nr = 0;
    int64_t checks = 0;
    zdx_pair_out out{};
    zdx_status st = ZDX_CAPACITY;
    int64_t loop_cap = 64, vtx_cap = 4096;
    for (int attempt = 0; attempt < 8 && st == ZDX_CAPACITY; ++attempt) {
        rec.assign((size_t)loop_cap * ZDX_LOOP_REC, 0.0);
        vtx.assign(3 * (size_t)vtx_cap, 0.0);
        vo.assign((size_t)loop_cap + 1, 0);
        out = zdx_pair_out{};
        out.tri_checks = &checks; out.n_segments = &ns; out.n_comps = &nc; out.n_loops_raw = &nr;
        out.loop_offset = lo.data(); out.loop_rec = rec.data(); out.loop_cap = loop_cap;
        out.vtx_offset = vo.data(); out.loop_vtx = vtx.data(); out.vtx_cap = vtx_cap;
        st = zdx_contact_pairs(ctx, &in, 1, &o, &out);
        if (st == ZDX_CAPACITY) {
            loop_cap = out.n_loops_total + 1;
            vtx_cap = out.n_vtx_total + 1;
        }
    }
    if (st != ZDX_OK) {
        std::fprintf(stderr, "zdx-pair: %s\n", zdx_last_error(ctx));
        zdx_destroy(ctx);
        return 1;
    }
    if (ns == 0) {
        std::printf("No intersection segments found.\n");
        zdx_destroy(ctx);
        return 0;
    }
    std::printf("Raw intersection segments: %d\n", ns);
    if (a.split) std::printf("Coarse components: %d (tol=%s)\n", nc, py_float_repr(a.tol).c_str());
    else std::printf("Coarse components: 1 (no split)\n");

    // polyline.py:780-834: the loops of each component in order, each one a
    // contact with its own id; the report numbers are polyline.py's, computed
    // from the loop's points
    const int nl = (int)out.n_loops_total;
    std::vector<std::vector<V3>> loops;
    std::vector<V3> xc0s;
    std::vector<char> has_xc0;
    std::vector<int> ids;
    std::vector<zdx::PolylineContact> pc;
    const int ncomp = a.split ? nc : 1;
    for (int comp = 0, l = 0; comp < ncomp; ++comp) {
        int e = l;
        while (e < nl && (int)rec[(size_t)e * ZDX_LOOP_REC + 23] == comp) ++e;
        std::vector<int> keep;   // polyline.py:798-806: at least 3 points, bbox diagonal >= 2 tol
        for (int m = l; m < e; ++m)
            if (zdx::polyline_keeps_loop(&vtx[3 * (size_t)vo[m]], (int)rec[(size_t)m * ZDX_LOOP_REC + 24], a.tol))
                keep.push_back(m);
        std::printf("\n[Component %d] loops=%d\n", comp, (int)keep.size());
        for (int m : keep) {
            const int nv = (int)rec[(size_t)m * ZDX_LOOP_REC + 24], cid = (int)pc.size();
            const double* p = &vtx[3 * (size_t)vo[m]];
            const zdx::PolylineContact c = zdx::polyline_contact(p, nv, a.k);
            std::vector<V3> pts;
            for (int v = 0; v < nv; ++v) pts.push_back(zdx::v3(p[3 * v], p[3 * v + 1], p[3 * v + 2]));
            std::printf("  [Contact %d] nVert=%d |Sn|=%s\n", cid, nv, py_float_repr(c.area).c_str());
            std::printf("    nA=%s\n", vec(c.nA).c_str());
            std::printf("    xc(0)=%s\n", c.has_xc0 ? vec(c.xc0).c_str() : "None");
            loops.push_back(pts);
            xc0s.push_back(zdx::v3(c.xc0[0], c.xc0[1], c.xc0[2]));
            has_xc0.push_back(c.has_xc0 ? 1 : 0);
            ids.push_back(cid);
            pc.push_back(c);
        }
        l = e;
    }
    if (!zdx::write_vtk_polydata_polylines(a.vtk, loops, xc0s, has_xc0, ids)) {
        std::fprintf(stderr, "zdx-pair: cannot write %s\n", a.vtk.c_str());
        zdx_destroy(ctx);
        return 1;
    }
    std::printf("\nReconstructed-loop VTK written: %s\n", a.vtk.c_str());
    std::printf("ParaView: color by CELL \"contact_id\". Each loop is one polyline cell.\n");
    std::printf("\n=== DEM ContactManifold ===\n");
    std::printf("idA=0 idB=1 nContact=%d\n", (int)pc.size());
    for (size_t l = 0; l < pc.size(); ++l) {
        const zdx::PolylineContact& c = pc[l];
        std::printf("  [cid=%d] nVert=%d area=%s Fn=%s\n", ids[l], (int)loops[l].size(), py_float_repr(c.area).c_str(),
                    py_float_repr(c.Fn).c_str());
        std::printf("    nA=%s xc0=%s F_A=%s\n", vec(c.nA).c_str(), c.has_xc0 ? vec(c.xc0).c_str() : "None",
                    vec(c.FA).c_str());
    }
    zdx_destroy(ctx);
    return 0;
}
