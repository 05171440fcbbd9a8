// zdx_api.cpp — the C ABI of libzdx.so (include/zdx.h): contexts, mesh
// registry, the batched per-pair contact call and the simulation step in its
// plain (one launch per stage, host-sized) and batched (persistent queue,
// hipGraph-captured step) forms.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/zdx.h"
#include "zdx_geom.h"
#include "zdx_internal.h"
#include "zdx_kernels.h"
#include "zdx_topo.h"

using namespace zdx;

namespace {

struct HipError {
    hipError_t e;
    const char* what;
};
#define HIPC(x)                                                \
    do {                                                       \
        hipError_t e_ = (x);                                   \
        if (e_ != hipSuccess) throw HipError{e_, #x};          \
    } while (0)

struct ApiError {
    zdx_status st;
    std::string msg;
};

// A fixed block of device memory handed out front to back (256-byte
// aligned pieces); nothing is returned before reset().
struct Arena {
    char* base = nullptr;
    size_t cap = 0, top = 0;
    void* carve(size_t bytes) {
        const size_t at = (top + 255) & ~(size_t)255;
        if (at + bytes > cap)
            throw ApiError{ZDX_CAPACITY, "the no-allocation scratch (" + std::to_string(cap >> 20) +
                                             " MiB) is too small for this call: use zdx_contact_pairs"};
        top = at + bytes;
        return base + at;
    }
    void reset() { top = 0; }
};

// device buffer that grows on demand (contents are not preserved on growth);
// bound to an Arena (ar != NULL) it takes its memory from there instead of
// hipMalloc and gives nothing back (release() only forgets the piece)
template <class T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0;
    Arena* ar = nullptr;
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf() { release(); }
    void release() {
        if (p && !ar) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    T* ensure(size_t m) {
        if (m == 0) m = 1;
        if (m <= n) return p;
        release();
        if (ar) p = (T*)ar->carve(m * sizeof(T));
        else HIPC(hipMalloc((void**)&p, m * sizeof(T)));
        n = m;
        return p;
    }
    void upload(const T* h, size_t m, hipStream_t s) {
        ensure(m);
        if (m) HIPC(hipMemcpyAsync(p, h, m * sizeof(T), hipMemcpyHostToDevice, s));
    }
};

struct MeshRec {
    HostMesh h;
    double tol, area_eps;
};

// Device buffers of the phased topology (zdx_topo.h), used by the time step
// and by the per-pair API: per arena segment (cells, components, vertices,
// edges), per work-list item (offsets, component counts), the bump-allocated
// adjacency / snap / T-junction buffers and the walks' global scratch slices.
struct TopoWork {
    DBuf<int32_t> cell9, comp, vcell, E, ncomp, lab;
    DBuf<double> V;
    DBuf<int32_t> adj;
    DBuf<uint64_t> snapw, tjw;
    DBuf<unsigned long long> top;
    DBuf<int32_t> units;
    DBuf<int64_t> adj_off, snap_off, tj_off;
    DBuf<int32_t> scr_i;
    DBuf<double> scr_d;
    int64_t adj_cap = 1 << 20, snap_cap = 1 << 17, tj_cap = 1 << 17, units_cap = 1 << 12;
    int32_t adj_row = 16;
    int64_t scr_ni = 0, scr_nd = 0;
    int64_t slices = 0;
    int32_t slice_seg_max = 0;

    void ensure_arena(int64_t segs) {
        cell9.ensure(9 * (size_t)segs); comp.ensure(TPC_N * (size_t)segs); V.ensure(6 * (size_t)segs);
        vcell.ensure(6 * (size_t)segs); E.ensure(2 * (size_t)segs); lab.ensure(2 * (size_t)segs);
    }
    void ensure_items(int64_t items) {
        adj_off.ensure(items); snap_off.ensure(items); tj_off.ensure(items); ncomp.ensure(items);
    }
    void ensure_slices(int64_t nslices, int32_t seg_max) {
        slices = nslices;
        slice_seg_max = seg_max;
        scr_ni = tp_walk_ints(seg_max, adj_row);
        scr_nd = tp_walk_dbls(seg_max);
        scr_i.ensure((size_t)(scr_ni * slices));
        scr_d.ensure((size_t)(scr_nd * slices));
    }
    void ensure_bump() {
        adj.ensure((size_t)adj_cap); snapw.ensure((size_t)snap_cap); tjw.ensure((size_t)tj_cap); top.ensure(5);
        units.ensure(2 * (size_t)units_cap);
    }
    // After a step whose pass A or walk B ran out of room: the bump counters
    // hold what every pair asked for.  Returns true if a buffer was grown.
    bool grow_from_counters(hipStream_t st) {
        unsigned long long t[5] = {0, 0, 0, 0, 0};
        HIPC(hipMemcpyAsync(t, top.p, sizeof(t), hipMemcpyDeviceToHost, st));
        HIPC(hipStreamSynchronize(st));
        bool g = false;
        auto grow = [&](int64_t& cap, unsigned long long need) {
            if ((int64_t)need > cap) { cap = (int64_t)need + (int64_t)need / 4 + 1024; g = true; }
        };
        if ((int64_t)t[3] > adj_row) {
            // longer rows: the rows buffer and the walk slices scale with them
            const int32_t r = (int32_t)t[3] + 4;
            t[0] = t[0] / (unsigned long long)(adj_row + 1) * (unsigned long long)(r + 1);
            adj_row = r;
            scr_i.release(); scr_d.release();
            ensure_slices(slices, slice_seg_max);
            g = true;
        }
        grow(adj_cap, t[0]);
        grow(snap_cap, t[1]);
        grow(tj_cap, t[2]);
        grow(units_cap, t[4]);
        if (g) {
            adj.release(); snapw.release(); tjw.release(); units.release();
            ensure_bump();
        }
        return g;
    }
    template <class F>
    void each_buf(F&& f) {
        f(cell9); f(comp); f(vcell); f(E); f(ncomp); f(lab); f(V); f(adj); f(snapw); f(tjw); f(top); f(units);
        f(adj_off); f(snap_off); f(tj_off); f(scr_i); f(scr_d);
    }
    void fill(TopoBufs& T) {
        ensure_bump();
        T.cell9 = cell9.p; T.adj = adj.p; T.adj_cap = adj_cap; T.adj_row = adj_row; T.snapw = snapw.p;
        T.snap_cap = snap_cap;
        T.tjw = tjw.p; T.tj_cap = tj_cap; T.top = top.p; T.adj_off = adj_off.p; T.snap_off = snap_off.p;
        T.units = units.p; T.units_cap = units_cap;
        T.tj_off = tj_off.p; T.ncomp = ncomp.p; T.comp = comp.p; T.V = V.p; T.vcell = vcell.p; T.E = E.p;
        T.lab = lab.p;
        T.scr_i = scr_i.p; T.scr_d = scr_d.p; T.scr_ni = scr_ni; T.scr_nd = scr_nd;
    }
};

}  // namespace

struct zdx_ctx {
    int device = 0;
    int n_cu = 256;
    hipStream_t stream = nullptr;
    // a second stream (default priority) on which the wall contacts run
    // alongside the narrow phase and the topology; fork/join events
    hipStream_t stream2 = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    // third stream: pass A of the many-segment pairs beside that of the others
    hipStream_t stream3 = nullptr;
    hipEvent_t ev_fork3 = nullptr, ev_join3 = nullptr, ev_cc3 = nullptr;
    // timing events of the per-pair call: transform, narrow phase (the
    // attempt that fitted), topology -- start and end of each
    hipEvent_t ev_api[6] = {};
    double api_ms[3] = {0.0, 0.0, 0.0};
    int32_t api_attempts = 0;
    std::string err;
    std::vector<MeshRec> meshes;
    bool mesh_dirty = true;
    uint64_t mesh_version = 0;
    int32_t max_tris = 1;
    // device mesh store
    DBuf<double> m_verts;
    DBuf<int32_t> m_tri;
    DBuf<int32_t> m_voff, m_nv, m_toff, m_nt;
    DBuf<double> m_tol, m_aeps;
    DBuf<int32_t> m_soff, m_ns, m_susp;
    // per-pair API buffers (api: grown with hipMalloc; api_na: carved from
    // na_arena, this device's no-allocation scratch, at every call)
    struct Api {
        DBuf<zdx_pair_in> pairs_in;
        DBuf<int32_t> mesh;
        DBuf<double> state;
        DBuf<int64_t> wv_off, wt_off;
        DBuf<double> wv, aabb;
        DBuf<int32_t> nondeg;
        DBuf<int32_t> pi, pj, order, npairs, status;
        DBuf<NarrowItem> items;
        DBuf<int64_t> seg_off, tri_checks, tri_tests;
        DBuf<int32_t> nseg, tri_hits, tri_scanned;
        DBuf<double> arena;
        DBuf<int32_t> list_i;
        DBuf<float> list_f;
        DBuf<int32_t> hits, hit_tri, hit_flag;
        DBuf<unsigned long long> hit_top;
        DBuf<int64_t> hit_off;
        DBuf<unsigned int> queue;
        DBuf<int32_t> heavy;
        DBuf<int32_t> bcnt, bcur;
        DBuf<int32_t> ncontrib, comps, loops_raw, tstatus, topo_list, topo_n, status2;
        TopoWork tw;
        DBuf<double> rec, vtx;
        DBuf<int32_t> prov;            // sensitivity runs
        DBuf<double> ej, vj, ljac;
        // packed output (device side of the copies to the caller)
        DBuf<int32_t> pk_loop_off, pk_seg_off, pk_vtx_off;
        DBuf<int64_t> pk_pvtx, pk_tot;
        DBuf<double> pk_rec, pk_ljac, pk_vtx, pk_seg;
        int32_t stage_cap = 1024;      // most segments one pair may store
        int64_t hit_cap = 1 << 16;     // hit slots = segment slots
        int32_t lmax = 32;

        template <class F>
        void each_buf(F&& f) {
            f(pairs_in); f(mesh); f(state); f(wv_off); f(wt_off); f(wv); f(aabb); f(nondeg); f(pi); f(pj); f(order);
            f(npairs); f(status); f(items); f(seg_off); f(tri_checks); f(tri_tests); f(nseg); f(tri_hits);
            f(tri_scanned); f(arena); f(list_i); f(list_f); f(hits); f(hit_tri); f(hit_flag); f(hit_top); f(hit_off);
            f(queue); f(heavy); f(bcnt); f(bcur); f(ncontrib); f(comps); f(loops_raw); f(tstatus); f(topo_list); f(topo_n);
            f(status2); f(rec); f(vtx); f(prov); f(ej); f(vj); f(ljac); f(pk_loop_off); f(pk_seg_off);
            f(pk_vtx_off); f(pk_pvtx); f(pk_tot); f(pk_rec); f(pk_ljac); f(pk_vtx); f(pk_seg);
            tw.each_buf(f);
        }
        // takes every buffer from `ar` from now on, starting from small
        // capacities (a call grows them inside the arena as needed)
        void bind(Arena* ar) {
            each_buf([&](auto& b) { b.release(); b.ar = ar; });
            stage_cap = 256; hit_cap = 1 << 15; lmax = 8;
            tw.adj_cap = 1 << 16; tw.snap_cap = 1 << 14; tw.tj_cap = 1 << 14; tw.units_cap = 1 << 10;
            tw.adj_row = 16; tw.slices = 0; tw.slice_seg_max = 0;
        }
    } api, api_na;
    // wall-contact queries (zdx_wall_contacts)
    struct WallQ {
        DBuf<int32_t> mesh, nondeg, isrc, status;
        DBuf<double> state, wv, aabb, tris, normals, scr, out;
        DBuf<int64_t> wv_off, wt_off;
    } wq;
    void* scratch_dev = nullptr;   // this device's `scratch` array
    Arena na_arena;

    DMeshes dmeshes() {
        DMeshes M;
        M.verts = m_verts.p; M.tri_idx = m_tri.p; M.vert_off = m_voff.p; M.nverts = m_nv.p;
        M.tri_off = m_toff.p; M.ntris = m_nt.p; M.tol = m_tol.p; M.area_eps = m_aeps.p;
        M.susp_off = m_soff.p; M.nsusp = m_ns.p; M.susp = m_susp.p;
        return M;
    }
    void sync_meshes() {
        if (!mesh_dirty) return;
        std::vector<double> v;
        std::vector<int32_t> t, voff, nv, toff, nt;
        std::vector<double> tol, aeps;
        std::vector<int32_t> soff, ns, susp;
        // meshes with the same triangle indices (scaled copies of one shape)
        // share one copy on the device, so that the narrow phase's index
        // loads stay in cache
        std::unordered_map<uint64_t, std::vector<int32_t>> seen;   // hash -> first triangles of index arrays
        for (auto& m : meshes) {
            soff.push_back((int32_t)susp.size());
            ns.push_back((int32_t)m.h.suspect.size());
            susp.insert(susp.end(), m.h.suspect.begin(), m.h.suspect.end());
            voff.push_back((int32_t)(v.size() / 3));
            nv.push_back((int32_t)(m.h.verts.size() / 3));
            nt.push_back((int32_t)(m.h.tri_idx.size() / 3));
            v.insert(v.end(), m.h.verts.begin(), m.h.verts.end());
            uint64_t hsh = 1469598103934665603ull ^ m.h.tri_idx.size();
            for (int32_t x : m.h.tri_idx) hsh = (hsh ^ (uint32_t)x) * 1099511628211ull;
            int32_t at = -1;
            for (int32_t o : seen[hsh])
                if (std::equal(m.h.tri_idx.begin(), m.h.tri_idx.end(), t.begin() + 3 * (size_t)o)) { at = o; break; }
            if (at < 0) {
                at = (int32_t)(t.size() / 3);
                seen[hsh].push_back(at);
                t.insert(t.end(), m.h.tri_idx.begin(), m.h.tri_idx.end());
            }
            toff.push_back(at);
            tol.push_back(m.tol);
            aeps.push_back(m.area_eps);
        }
        m_verts.upload(v.data(), v.size(), stream);
        m_tri.upload(t.data(), t.size(), stream);
        m_voff.upload(voff.data(), voff.size(), stream);
        m_nv.upload(nv.data(), nv.size(), stream);
        m_toff.upload(toff.data(), toff.size(), stream);
        m_nt.upload(nt.data(), nt.size(), stream);
        m_tol.upload(tol.data(), tol.size(), stream);
        m_aeps.upload(aeps.data(), aeps.size(), stream);
        m_soff.upload(soff.data(), soff.size(), stream);
        m_ns.upload(ns.data(), ns.size(), stream);
        m_susp.upload(susp.data(), susp.size(), stream);
        HIPC(hipStreamSynchronize(stream));
        mesh_dirty = false;
        ++mesh_version;
    }
    int add_mesh(HostMesh&& h) {
        MeshRec r;
        // tol and area_eps of mesh A (main.cpp:2118-2121, 2205)
        double tol = h.mean_edge > 0.0 ? (h.mean_edge * 0.1) : (h.bbox_diag * 1e-2);
        if (tol <= 0.0) tol = 1e-6;
        r.tol = tol;
        r.area_eps = h.mean_edge > 0.0 ? (h.mean_edge * h.mean_edge * 1e-4) : 1e-12;
        max_tris = std::max<int32_t>(max_tris, (int32_t)(h.tri_idx.size() / 3));
        r.h = std::move(h);
        meshes.push_back(std::move(r));
        mesh_dirty = true;
        return (int)meshes.size() - 1;
    }
};

namespace {

template <class F>
zdx_status guard(zdx_ctx* ctx, F&& f) {
    try {
        f();
        return ZDX_OK;
    } catch (const HipError& e) {
        if (ctx) ctx->err = std::string("HIP error ") + hipGetErrorString(e.e) + " in " + e.what;
        return ZDX_HIP_ERROR;
    } catch (const ApiError& e) {
        if (ctx) ctx->err = e.msg;
        return e.st;
    } catch (const std::exception& e) {
        if (ctx) ctx->err = e.what();
        return ZDX_INTERNAL;
    }
}

// Device memory of zdx_contact_pairs_noalloc: one array per GPU, made by the
// code object's loader, so that entry point calls no allocator.  Host code
// never uses the name `scratch` as an address: each context looks up this
// device's copy with hipGetSymbolAddress when it is created.
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)256 << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];
// bounds of the no-allocation call's workgroup-sized buffers: the cull's
// per-workgroup triangle lists (80 bytes per triangle of the largest mesh)
// and the topology walks' scratch slices
constexpr int64_t NOALLOC_LIST_BYTES = (int64_t)16 << 20;
constexpr int64_t NOALLOC_WALK_SLICES = 256;

struct CallLimits {
    bool may_allocate = true;                       // false: device memory only from the arena
    int64_t list_bytes = (int64_t)1 << 60;          // cap of the cull's per-workgroup lists
    int64_t walk_slices = 4096;                     // most topology walk workgroups
};

// ZDX_START_CAPACITY=tiny: every growable buffer of a context or simulation
// created while it is set starts at a few entries, so the first calls and
// steps run the grow-and-retry paths (results do not depend on it).
bool tiny_start_capacity() {
    const char* e = std::getenv("ZDX_START_CAPACITY");
    return e && std::strcmp(e, "tiny") == 0;
}

int32_t persistent_groups(const zdx_ctx* c, int per_cu) { return std::max(1, c->n_cu * per_cu); }

// The hit arena after an overflow: each shard's counter holds what its pairs
// asked for, so twice the largest one sizes every shard (the shards' loads
// differ and change from step to step; the arena at least doubles).
int64_t grown_hit_cap(const unsigned long long* d_top, int shards, int64_t cap) {
    std::vector<unsigned long long> top((size_t)shards * HIT_TOP_STRIDE, 0);
    HIPC(hipMemcpy(top.data(), d_top, top.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    unsigned long long want = 0;
    for (int k = 0; k < shards; ++k) want = std::max(want, top[(size_t)k * HIT_TOP_STRIDE]);
    return std::max<int64_t>(2 * cap, (int64_t)shards * (int64_t)(2 * want + 1));
}

}  // namespace

// ======================================================================= sim
struct zdx_sim {
    zdx_ctx* ctx = nullptr;
    int32_t N = 0, W = 0;
    SimConfig cfg;
    std::vector<ParticleInit> inits;   // per particle, as configured (state txt output)
    SimParams sp{};
    std::vector<ParticleConst> pc;
    std::vector<int32_t> mesh_of;
    int64_t n_steps_done = 0;
    int parity = 0;
    // per-particle constants
    DBuf<int32_t> d_mesh;
    DBuf<double> d_mass, d_inv_mass, d_Ibinv, d_eqr, d_young, d_poisson, d_mu, d_rest, d_radius;
    DBuf<int64_t> d_wv_off, d_wt_off;
    DBuf<double> d_wv, d_aabb;
    DBuf<int32_t> d_nondeg;
    DBuf<int32_t> d_need_list, d_n_need;
    DBuf<int32_t> d_wall_list, d_n_wall;
    DBuf<double> d_state[2];
    DBuf<double> d_F, d_T;
    DBuf<int32_t> d_counts;
    // broad phase
    DBuf<DGrid> d_grid;
    DBuf<double> d_bp_part, d_spos;
    int32_t max_cells = 1 << 24;
    DBuf<int32_t> d_cell_of, d_cell_count, d_cell_start, d_sorted, d_cnt_up, d_cnt_all, d_nbr;
    DBuf<int32_t> d_pair_off, d_adj_off, d_adj_j, d_adj_pair, d_n_adj, d_scan_tmp;
    int32_t pair_cap = 0, adj_cap = 0;
    // pair sets (double-buffered: this step / previous step for the histories)
    struct PairSet {
        DBuf<int32_t> pi, pj, n;
        DBuf<int64_t> key;
        DBuf<double> ds;
        DBuf<uint8_t> active;
        DBuf<double> wds;
        DBuf<uint8_t> wactive;
    } ps[2];
    DBuf<int32_t> d_order, d_bcnt, d_bcur, d_topo_list, d_topo_n, d_topo_big, d_topo_small, d_topo_nbs;
    DBuf<NarrowItem> d_items;
    DBuf<int64_t> d_seg_off, d_tri_checks, d_tri_tests;
    DBuf<int32_t> d_nseg, d_tri_hits, d_tri_scanned;
    DBuf<double> d_arena;        // segments, one slot per hit slot
    DBuf<int32_t> d_list_i;
    DBuf<float> d_list_f;
    DBuf<int32_t> d_hits, d_hit_flag;
    DBuf<unsigned long long> d_hit_top;
    DBuf<int64_t> d_hit_off;
    int64_t hit_cap = 0;
    int32_t stage_cap = 512;   // most segments one pair may store
    int32_t tmax = 1;            // largest triangle count of the meshes in use
    uint64_t mesh_version = 0;   // context mesh upload the graphs were captured with
    DBuf<unsigned int> d_queue;
    DBuf<int32_t> d_heavy;
    // topology
    DBuf<double> d_contrib;
    DBuf<int32_t> d_ncontrib, d_comps, d_loops_raw, d_tstatus;
    TopoWork tw;
    int32_t lmax = 8;
    int64_t topo_threads = 2048;   // persistent one-wave workgroups of the topology walks
    int64_t topo_pass_groups = 1024;  // persistent 256-thread workgroups of the topology passes
    int32_t topo_lds = 32 * 1024;  // LDS scratch per walk workgroup (pairs up to ~120 segments)
    int32_t seg_max = 512;
    // walls
    DBuf<double> d_wtris, d_wnorm, d_wpos, d_wrad, d_wmu;
    DBuf<int32_t> d_wtoff, d_wnt;
    int32_t max_wall_tris = 1;
    DBuf<double> d_wcontrib;
    DBuf<int32_t> d_wnc;
    DBuf<double> d_wscr;
    int64_t wall_threads = 8192;
    int32_t wall_pts_cap = 64;
    // status / counters
    DBuf<int32_t> d_status;
    DBuf<unsigned long long> d_ctr;
    int64_t last_ctr[ZDX_N_COUNTERS] = {0};
    double stage_ms[6] = {0, 0, 0, 0, 0, 0};
    NarrowCfg ncfg{256, 4, 0, 0};
    int32_t np_groups_per_cu = 8;
    // per-stage kernel spans stamped on the device clock inside every step
    // (also inside the captured graphs): 6 stages x (start, end)
    DBuf<unsigned long long> d_kt;
    double clock_khz = 100000.0;
    double acc_ms[7] = {0, 0, 0, 0, 0, 0, 0};   // 6 stages, then the narrow-phase kernels alone
    double acc_ctr[ZDX_N_COUNTERS] = {0};
    int64_t acc_steps = 0;
    // steps run again after a capacity growth, and graph captures, since the
    // last zdx_sim_stage_times call
    int64_t acc_retries = 0, acc_captures = 0;
    // graphs (one per parity)
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    hipGraph_t graph[2] = {nullptr, nullptr};
    bool graphs_valid = false;

    // end-of-step report (pinned, coherent host memory written by k_report)
    StepReport* rep = nullptr;

    ~zdx_sim() {
        drop_graphs();
        if (rep) (void)hipHostFree(rep);
    }
    void drop_graphs() {
        for (int k = 0; k < 2; ++k) {
            if (gexec[k]) (void)hipGraphExecDestroy(gexec[k]);
            if (graph[k]) (void)hipGraphDestroy(graph[k]);
            gexec[k] = nullptr;
            graph[k] = nullptr;
        }
        graphs_valid = false;
    }

    hipStream_t st() const { return ctx->stream; }

    DBodies bodies(int par) {
        DBodies B;
        B.n = N; B.mesh = d_mesh.p; B.state = d_state[par].p; B.wv_off = d_wv_off.p; B.wt_off = d_wt_off.p;
        B.wv = d_wv.p; B.nondeg = d_nondeg.p; B.aabb = d_aabb.p;
        return B;
    }
    DMat mat() {
        DMat m;
        m.mass = d_mass.p; m.inv_mass = d_inv_mass.p; m.Ibinv = d_Ibinv.p; m.eqr = d_eqr.p; m.young = d_young.p;
        m.poisson = d_poisson.p; m.mu = d_mu.p; m.rest = d_rest.p; m.radius = d_radius.p;
        return m;
    }
    BroadPhaseBufs bpbufs(int par) {
        BroadPhaseBufs b;
        b.n = N; b.state = d_state[par].p; b.radius = d_radius.p; b.grid = d_grid.p; b.max_cells = max_cells;
        b.part = d_bp_part.p; b.spos = d_spos.p;
        b.cell_of = d_cell_of.p; b.cell_count = d_cell_count.p; b.cell_start = d_cell_start.p;
        b.sorted = d_sorted.p; b.cnt_up = d_cnt_up.p; b.cnt_all = d_cnt_all.p; b.nbr = d_nbr.p;
        b.pair_off = d_pair_off.p; b.adj_off = d_adj_off.p; b.pi = ps[par].pi.p; b.pj = ps[par].pj.p;
        b.pkey = ps[par].key.p; b.pair_cap = pair_cap; b.adj_j = d_adj_j.p; b.adj_pair = d_adj_pair.p;
        b.adj_cap = adj_cap; b.n_pairs = ps[par].n.p; b.n_adj = d_n_adj.p; b.scan_tmp = d_scan_tmp.p;
        b.status = d_status.p;
        return b;
    }
    DPairs dpairs(int par, bool lpt) {
        DPairs P{};
        P.cap = pair_cap; P.n = ps[par].n.p; P.pi = ps[par].pi.p; P.pj = ps[par].pj.p;
        P.order = lpt ? d_order.p : nullptr;
        P.items = lpt ? d_items.p : nullptr;
        P.seg_off = d_seg_off.p; P.nseg = d_nseg.p; P.tri_checks = d_tri_checks.p; P.tri_tests = d_tri_tests.p;
        P.tri_hits = d_tri_hits.p; P.tri_scanned = d_tri_scanned.p; P.arena = d_arena.p; P.arena_cap = hit_cap;
        P.status = d_status.p;
        P.topo_list = d_topo_list.p; P.topo_n = d_topo_n.p;
        P.topo_big = d_topo_big.p; P.topo_small = d_topo_small.p; P.n_big = d_topo_nbs.p; P.n_small = d_topo_nbs.p + 1;
        P.hits = d_hits.p; P.hit_cap = hit_cap; P.hit_top = d_hit_top.p; P.hit_off = d_hit_off.p;
        P.hit_shards = (int32_t)std::min<int64_t>(HIT_SHARDS, hit_cap);
        P.hit_flag = d_hit_flag.p;
        P.heavy = d_heavy.p;
        return P;
    }
    HistBufs hist(int par) {
        HistBufs H;
        const int pv = 1 - par;
        H.prev_key = ps[pv].key.p; H.prev_ds = ps[pv].ds.p; H.prev_active = ps[pv].active.p; H.prev_n = ps[pv].n.p;
        H.cur_ds = ps[par].ds.p; H.cur_active = ps[par].active.p;
        H.wprev_ds = ps[pv].wds.p; H.wprev_active = ps[pv].wactive.p;
        H.wcur_ds = ps[par].wds.p; H.wcur_active = ps[par].wactive.p;
        return H;
    }
    TopoBufs topobufs() {
        TopoBufs T{};
        T.lmax = lmax; T.contrib = d_contrib.p; T.ncontrib = d_ncontrib.p; T.comps = d_comps.p;
        T.loops_raw = d_loops_raw.p; T.tstatus = d_tstatus.p;
        tw.fill(T);
        T.nthreads = topo_threads; T.pass_groups = topo_pass_groups; T.seg_max = seg_max; T.lds_bytes = topo_lds;
        return T;
    }
    WallBufs wallbufs() {
        WallBufs Wb;
        Wb.n_walls = W; Wb.tris = d_wtris.p; Wb.normals = d_wnorm.p; Wb.tri_off = d_wtoff.p; Wb.ntri = d_wnt.p;
        Wb.pos = d_wpos.p; Wb.radius = d_wrad.p; Wb.mu = d_wmu.p; Wb.max_wall_tris = max_wall_tris;
        return Wb;
    }

    void alloc_pairs(int32_t cap) {
        drop_graphs();
        pair_cap = cap;
        adj_cap = 2 * cap;
        for (int k = 0; k < 2; ++k) {
            // preserve the previous step's history when growing
            std::vector<int64_t> key;
            std::vector<double> ds;
            std::vector<uint8_t> act;
            int32_t n = 0;
            if (ps[k].n.p) {
                HIPC(hipMemcpy(&n, ps[k].n.p, sizeof(int32_t), hipMemcpyDeviceToHost));
                if (n > 0) {
                    key.resize(n); ds.resize(3 * (size_t)n); act.resize(n);
                    HIPC(hipMemcpy(key.data(), ps[k].key.p, n * sizeof(int64_t), hipMemcpyDeviceToHost));
                    HIPC(hipMemcpy(ds.data(), ps[k].ds.p, 3 * n * sizeof(double), hipMemcpyDeviceToHost));
                    HIPC(hipMemcpy(act.data(), ps[k].active.p, n * sizeof(uint8_t), hipMemcpyDeviceToHost));
                }
            }
            ps[k].pi.ensure(cap); ps[k].pj.ensure(cap); ps[k].n.ensure(1);
            if (ps[k].key.n < (size_t)cap) {
                ps[k].key.release(); ps[k].ds.release(); ps[k].active.release();
            }
            ps[k].key.ensure(cap); ps[k].ds.ensure(3 * (size_t)cap); ps[k].active.ensure(cap);
            if (n > 0) {
                HIPC(hipMemcpy(ps[k].key.p, key.data(), n * sizeof(int64_t), hipMemcpyHostToDevice));
                HIPC(hipMemcpy(ps[k].ds.p, ds.data(), 3 * n * sizeof(double), hipMemcpyHostToDevice));
                HIPC(hipMemcpy(ps[k].active.p, act.data(), n * sizeof(uint8_t), hipMemcpyHostToDevice));
            }
            HIPC(hipMemcpy(ps[k].n.p, &n, sizeof(int32_t), hipMemcpyHostToDevice));
        }
        d_adj_j.ensure(adj_cap); d_adj_pair.ensure(adj_cap);
        d_order.ensure(cap); d_items.ensure(cap); d_heavy.ensure(cap); d_topo_list.ensure(cap); d_topo_n.ensure(1);
        d_topo_big.ensure(cap); d_topo_small.ensure(cap); d_topo_nbs.ensure(2);
        d_hit_off.ensure(cap); d_hit_top.ensure(HIT_SHARDS * HIT_TOP_STRIDE);
        d_seg_off.ensure(cap); d_tri_checks.ensure(cap); d_tri_tests.ensure(cap);
        d_nseg.ensure(cap); d_tri_hits.ensure(cap); d_tri_scanned.ensure(cap);
        d_ncontrib.ensure(cap); d_comps.ensure(cap); d_loops_raw.ensure(cap); d_tstatus.ensure(cap);
        d_contrib.ensure((size_t)cap * lmax * 9);
        tw.ensure_items(cap);
    }
    void alloc_lmax(int32_t l) {
        drop_graphs();
        lmax = l;
        d_contrib.release();
        d_contrib.ensure((size_t)pair_cap * lmax * 9);
    }
    void alloc_stage(int32_t cap) {
        drop_graphs();
        stage_cap = cap;
        seg_max = std::max(seg_max, cap);
        alloc_topo_scratch();
    }
    void alloc_hits(int64_t cap) {
        drop_graphs();
        hit_cap = cap;
        d_hits.release(); d_hit_flag.release(); d_arena.release();
        d_hits.ensure(6 * (size_t)cap); d_hit_flag.ensure((size_t)cap); d_arena.ensure(6 * (size_t)cap);
        tw.ensure_arena(cap);
    }
    void alloc_topo_scratch() {
        drop_graphs();
        tw.scr_i.release(); tw.scr_d.release();
        tw.ensure_slices(topo_threads, seg_max);
    }
    int32_t transform_groups() const { return std::max(1, ctx->n_cu * 8); }
};

namespace {

void sim_setup_device(zdx_sim* s) {
    zdx_ctx* c = s->ctx;
    c->sync_meshes();
    const int N = s->N;
    std::vector<double> mass(N), im(N), Ibinv(9 * (size_t)N), eqr(N), young(N), pois(N), mu(N), rest(N), rad(N);
    std::vector<int64_t> wvo(N + 1), wto(N + 1);
    int64_t av = 0, at = 0;
    for (int i = 0; i < N; ++i) {
        const ParticleConst& p = s->pc[i];
        mass[i] = p.mass; im[i] = p.inv_mass; eqr[i] = p.eqr; young[i] = p.young; pois[i] = p.poisson;
        mu[i] = p.mu; rest[i] = p.rest; rad[i] = p.radius;
        for (int k = 0; k < 9; ++k) Ibinv[9 * (size_t)i + k] = p.Ibinv[k];
        wvo[i] = av; wto[i] = at;
        av += (int64_t)(c->meshes[p.mesh].h.verts.size() / 3);
        at += (int64_t)(c->meshes[p.mesh].h.tri_idx.size() / 3);
    }
    wvo[N] = av; wto[N] = at;
    if (av >= ((int64_t)1 << 31))   // the narrow phase's hit records index world vertices with int32
        throw ApiError{ZDX_CAPACITY, "more than 2^31 world vertices in one simulation"};
    hipStream_t st = c->stream;
    s->d_mesh.upload(s->mesh_of.data(), N, st);
    s->d_mass.upload(mass.data(), N, st); s->d_inv_mass.upload(im.data(), N, st);
    s->d_Ibinv.upload(Ibinv.data(), Ibinv.size(), st); s->d_eqr.upload(eqr.data(), N, st);
    s->d_young.upload(young.data(), N, st); s->d_poisson.upload(pois.data(), N, st);
    s->d_mu.upload(mu.data(), N, st); s->d_rest.upload(rest.data(), N, st); s->d_radius.upload(rad.data(), N, st);
    s->d_wv_off.upload(wvo.data(), N + 1, st); s->d_wt_off.upload(wto.data(), N + 1, st);
    s->d_wv.ensure(3 * (size_t)av); s->d_aabb.ensure(6 * (size_t)N);
    s->d_nondeg.ensure(N);
    s->d_need_list.ensure(N); s->d_n_need.ensure(1);
    s->d_wall_list.ensure(N); s->d_n_wall.ensure(1);
    s->d_state[0].ensure(16 * (size_t)N); s->d_state[1].ensure(16 * (size_t)N);
    s->d_F.ensure(3 * (size_t)N); s->d_T.ensure(3 * (size_t)N); s->d_counts.ensure(N);
    // broad phase
    s->d_grid.ensure(1);
    s->d_bp_part.ensure(7 * (size_t)BP_PARTS);
    s->d_spos.ensure(4 * (size_t)N);
    int64_t cells = std::min<int64_t>((int64_t)1 << 24, std::max<int64_t>(64, 8 * (int64_t)N));
    s->max_cells = (int32_t)cells;
    s->d_cell_of.ensure(N); s->d_cell_count.ensure(cells); s->d_cell_start.ensure(cells + 1);
    s->d_sorted.ensure(N); s->d_nbr.ensure(16 * (size_t)N); s->d_cnt_up.ensure(N); s->d_cnt_all.ensure(N);
    s->d_pair_off.ensure(N + 1); s->d_adj_off.ensure(N + 1); s->d_n_adj.ensure(1);
    s->d_scan_tmp.ensure((size_t)(2 * (std::max<int64_t>(cells, N) / 1024 + 4)));
    s->d_status.ensure(1); s->d_ctr.ensure(16); s->d_queue.ensure(NARROW_COUNTERS);
    s->d_bcnt.ensure(64); s->d_bcur.ensure(64);
    s->d_kt.ensure(12);
    // the cell counters return to zero within every broad phase (k_bp_scatter)
    launch_fill_i32(s->d_cell_count.p, 0, cells, c->stream);
    HIPC(hipHostMalloc((void**)&s->rep, sizeof(StepReport), hipHostMallocCoherent));
    std::memset((void*)s->rep, 0, sizeof(StepReport));
    launch_fill_u64(s->d_ctr.p, 0ull, 16, c->stream);
    launch_fill_i32(s->d_status.p, 0, 1, c->stream);
    {
        int khz = 0;
        if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, c->device) == hipSuccess && khz > 0)
            s->clock_khz = (double)khz;
    }
    const bool tiny = tiny_start_capacity();
    if (tiny) {
        s->stage_cap = 4; s->seg_max = 4; s->lmax = 1;
        s->tw.adj_cap = 64; s->tw.snap_cap = 64; s->tw.tj_cap = 64; s->tw.units_cap = 4;
    }
    s->alloc_pairs(tiny ? 8 : std::max(1024, 8 * N));
    size_t groups = (size_t)persistent_groups(c, 16);
    s->tmax = c->max_tris;
    s->d_list_i.ensure(groups * 8 * (size_t)s->tmax);
    s->d_list_f.ensure(groups * 12 * (size_t)s->tmax);
    s->topo_threads = std::min<int64_t>(2048, std::max<int64_t>(64, (int64_t)N));
    s->alloc_stage(s->stage_cap);
    s->alloc_hits(tiny ? 64 : std::max<int64_t>(1 << 16, 32 * (int64_t)N));
    s->tw.ensure_bump();   // before any graph capture
    // particle/wall histories start empty
    for (int k = 0; k < 2; ++k) {
        int32_t z = 0;
        HIPC(hipMemcpyAsync(s->ps[k].n.p, &z, sizeof(int32_t), hipMemcpyHostToDevice, st));
    }
    HIPC(hipStreamSynchronize(st));
}

struct WallHost {
    std::vector<double> tris, normals;  // world
    double pos[3];
    double radius, mu, rest;
};

void sim_setup_walls(zdx_sim* s, const std::vector<WallHost>& walls) {
    hipStream_t st = s->st();
    s->W = (int32_t)walls.size();
    std::vector<double> wt, wn, wp, wr, wmu;
    std::vector<int32_t> wto, wnt;
    int32_t maxwt = 1;
    for (const auto& w : walls) {
        wto.push_back((int32_t)(wt.size() / 9));
        wnt.push_back((int32_t)(w.tris.size() / 9));
        maxwt = std::max<int32_t>(maxwt, (int32_t)(w.tris.size() / 9));
        wt.insert(wt.end(), w.tris.begin(), w.tris.end());
        wn.insert(wn.end(), w.normals.begin(), w.normals.end());
        wp.insert(wp.end(), w.pos, w.pos + 3);
        wr.push_back(w.radius);
        wmu.push_back(w.mu);
    }
    s->max_wall_tris = maxwt;
    const size_t npw = (size_t)s->N * std::max(1, s->W);
    s->d_wtris.upload(wt.data(), wt.size(), st); s->d_wnorm.upload(wn.data(), wn.size(), st);
    s->d_wpos.upload(wp.data(), wp.size(), st); s->d_wrad.upload(wr.data(), wr.size(), st);
    s->d_wmu.upload(wmu.data(), wmu.size(), st); s->d_wtoff.upload(wto.data(), wto.size(), st);
    s->d_wnt.upload(wnt.data(), wnt.size(), st);
    s->d_wcontrib.ensure(npw * (size_t)maxwt * 6);
    s->d_wnc.ensure(npw);
    launch_fill_i32(s->d_wnc.p, 0, (int64_t)npw, st);
    for (int k = 0; k < 2; ++k) {
        s->ps[k].wds.ensure(3 * npw);
        s->ps[k].wactive.ensure(npw);
        HIPC(hipMemsetAsync(s->ps[k].wactive.p, 0, npw, st));
    }
    s->wall_pts_cap = tiny_start_capacity() ? 4 : std::max<int32_t>(64, s->ctx->max_tris);
    s->wall_threads = std::min<int64_t>(2048, std::max<int64_t>(1, (int64_t)s->N));   // one-wave workgroups
    s->d_wscr.ensure((size_t)s->wall_threads * (size_t)(3 * s->wall_pts_cap + 4 * (s->wall_pts_cap + 8)));
    HIPC(hipStreamSynchronize(st));
}

WallHost make_wall(const HostMesh& m, const double* pos, const double* quat, double mu, double rest) {
    WallHost w;
    Qt q; q.w = quat[0]; q.x = quat[1]; q.y = quat[2]; q.z = quat[3];
    V3 p = v3(pos[0], pos[1], pos[2]);
    const size_t T = m.tris.size() / 9;
    for (size_t t = 0; t < T; ++t) {
        V3 c[3];
        for (int k = 0; k < 3; ++k) {
            const double* v = &m.tris[9 * t + 3 * k];
            c[k] = add(qrotate(q, v3(v[0], v[1], v[2])), p);  // transform_tris main.cpp:499-510
            w.tris.push_back(c[k].x); w.tris.push_back(c[k].y); w.tris.push_back(c[k].z);
        }
        V3 n = normalize(cross(sub(c[1], c[0]), sub(c[2], c[0])));  // tri_normal main.cpp:512-516
        w.normals.push_back(n.x); w.normals.push_back(n.y); w.normals.push_back(n.z);
    }
    std::memcpy(w.pos, pos, 3 * sizeof(double));
    w.radius = m.radius;
    w.mu = mu;
    w.rest = rest;
    return w;
}

// Launch sequence of one step.  Batched: everything device-sized (persistent
// grids reading counts from device memory, LPT work queue) so the step can be
// captured in a graph.  Plain: one launch per stage with the grids sized on
// the host from the counts read back after the broad phase and after the
// narrow phase.
void enqueue_step(zdx_sim* s, int par, bool batched) {
    hipStream_t st = s->st();
    const int nxt = 1 - par;
    DBodies B = s->bodies(par);
    DMeshes M = s->ctx->dmeshes();
    DMat mat = s->mat();
    // stage stamps: 0 start, 1 after the broad phase, 2 after the transform,
    // 3 after the narrow phase, 4 after the topology, 5 when the wall stream has
    // joined, 6 after fold + integrate, 7 after the walls (on the wall stream,
    // which starts at 2 and runs alongside 2..4), 8 before the narrow-phase
    // kernels (after the work ordering), 9 when the wall stream starts its
    // kernels.  Most are written by the first
    // workgroup of the kernel that starts the next stage; 7 (and 2, 3 in
    // plain mode, where the host reads counts in between) by a stamp kernel.
    unsigned long long* kt = s->d_kt.p;
    const int gmax = persistent_groups(s->ctx, s->np_groups_per_cu);
    {
        StepZero z{};
        z.i32[0] = s->d_status.p; z.n32[0] = 1;
        z.i32[1] = s->d_n_need.p; z.n32[1] = 1;
        z.i32[2] = s->d_n_wall.p; z.n32[2] = 1;
        z.i32[3] = s->d_topo_n.p; z.n32[3] = 1;
        z.i32[4] = (int32_t*)s->d_queue.p; z.n32[4] = NARROW_COUNTERS;
        z.i32[5] = batched ? s->d_bcnt.p : nullptr; z.n32[5] = 64;
        z.i32[7] = batched ? s->d_bcur.p : nullptr; z.n32[7] = 64;
        z.i32[6] = s->d_topo_nbs.p; z.n32[6] = 2;
        z.u64[0] = s->d_hit_top.p; z.n64[0] = HIT_SHARDS * HIT_TOP_STRIDE;
        z.u64[1] = s->tw.top.p; z.n64[1] = 5;
        // the step's start (clock 0 and z) is done by the broad phase's first
        // kernel; the need lists come from its count pass
        const NeedLists need{s->wallbufs(), s->d_need_list.p, s->d_n_need.p, s->d_wall_list.p, s->d_n_wall.p};
        launch_broadphase(s->bpbufs(par), st, &need, kt, &z);
    }
    int g_transform = s->transform_groups(), g_narrow = gmax, g_topo = (int)s->topo_threads,
        g_walls = (int)s->wall_threads;
    if (!batched) {
        int32_t c[3] = {0, 0, 0};
        HIPC(hipMemcpyAsync(&c[0], s->ps[par].n.p, 4, hipMemcpyDeviceToHost, st));
        HIPC(hipMemcpyAsync(&c[1], s->d_n_need.p, 4, hipMemcpyDeviceToHost, st));
        HIPC(hipMemcpyAsync(&c[2], s->d_n_wall.p, 4, hipMemcpyDeviceToHost, st));
        HIPC(hipStreamSynchronize(st));
        g_narrow = std::max(1, std::min(c[0], gmax));
        g_transform = std::max(1, std::min(c[1], g_transform));
        g_walls = std::max(1, std::min(c[2], g_walls));
    }
    launch_transform(B, M, st, s->d_need_list.p, s->d_n_need.p, g_transform, kt + 1);
    if (!batched) launch_stamp(kt + 2, st);   // batched: stamped by the work ordering's first kernel
    const bool walls = s->W > 0;
    if (walls) {
        // the wall contacts need only the world vertices and the wall list:
        // fork them onto the second stream beside the narrow phase and the
        // topology, join before the fold (forked after the narrow phase they
        // slow the topology's pass A more than they cost the narrow phase)
        hipStream_t st2 = s->ctx->stream2;
        HIPC(hipEventRecord(s->ctx->ev_fork, st));
        HIPC(hipStreamWaitEvent(st2, s->ctx->ev_fork, 0));
        launch_walls(B, M, mat, s->wallbufs(), s->hist(par), s->d_wcontrib.p, s->d_wnc.p, s->d_wall_list.p,
                     s->d_n_wall.p, s->d_wscr.p, g_walls, s->wall_pts_cap, s->sp, s->d_status.p, st2, kt + 9);
        launch_stamp(kt + 7, st2);
        HIPC(hipEventRecord(s->ctx->ev_join, st2));
    }
    DPairs P = s->dpairs(par, batched);
    TopoBufs T = s->topobufs();
    HistBufs H = s->hist(par);
    // batched: the work ordering's scatter pass zeroes the topology's per-pair
    // outputs, and k_tp_alloc stamps the narrow phase's end
    if (batched)
        launch_lpt_order(P, M, B, s->d_order.p, s->d_bcnt.p, s->d_bcur.p, s->pair_cap, st, true, kt + 2, &T,
                         H.cur_active);
    NarrowCfg nc = s->ncfg;
    nc.use_queue = batched ? 1 : 0;
    nc.groups = g_narrow;
    nc.wave_groups = persistent_groups(s->ctx, 4);
    launch_narrow(B, M, P, nc, s->d_list_i.p, s->d_list_f.p, s->stage_cap, s->tmax, s->d_queue.p, st, true, kt + 8);
    if (!batched) launch_stamp(kt + 3, st);   // batched: stamped by k_tp_alloc (no host read in between)
    if (!batched) {
        int32_t nt = 0;
        HIPC(hipMemcpyAsync(&nt, s->d_topo_n.p, 4, hipMemcpyDeviceToHost, st));
        HIPC(hipStreamSynchronize(st));
        T.nthreads = std::max(1, std::min<int>(nt, g_topo));
        T.pass_groups = std::max<int64_t>(1, std::min<int64_t>(nt, T.pass_groups));
    }
    if (batched) T.stamp_alloc = kt + 3;
    else launch_topo_zero(P, T, H.cur_active, s->pair_cap, st, nullptr);
    launch_topo_sim(P, M, B, mat, H, T, s->sp, st, true, s->ctx->stream3, s->ctx->ev_fork3, s->ctx->ev_join3,
                    s->ctx->ev_cc3);
    if (walls) HIPC(hipStreamWaitEvent(st, s->ctx->ev_join, 0));
    // stamps 4 (topology done, with any wait for the wall stream) and 5 (fold
    // start) come from the fold's first workgroup, 6 from the report's
    launch_fold_integrate(s->N, s->d_state[par].p, s->d_state[nxt].p, mat, s->d_adj_off.p, s->d_adj_j.p,
                          s->d_adj_pair.p, s->adj_cap, T, s->W, s->d_wcontrib.p, s->d_wnc.p, s->max_wall_tris,
                          s->d_F.p, s->d_T.p, s->d_counts.p, s->sp, P, s->d_ctr.p, st, kt + 4, kt + 5);
    launch_report(s->d_status.p, s->d_ctr.p, kt, s->rep, st);
}

int32_t read_status(zdx_sim* s) {
    int32_t v = 0;
    HIPC(hipMemcpyAsync(&v, s->d_status.p, sizeof(int32_t), hipMemcpyDeviceToHost, s->st()));
    HIPC(hipStreamSynchronize(s->st()));
    return v;
}

// Grow whatever a failed step reported; returns false when nothing can grow.
bool grow_for(zdx_sim* s, int32_t status, int par) {
    bool grown = false;
    if (status & ST_PAIRS) {
        int32_t np = 0;
        HIPC(hipMemcpy(&np, s->d_pair_off.p + s->N, sizeof(int32_t), hipMemcpyDeviceToHost));
        int32_t na = 0;
        HIPC(hipMemcpy(&na, s->d_adj_off.p + s->N, sizeof(int32_t), hipMemcpyDeviceToHost));
        int32_t need = std::max(np, (na + 1) / 2);
        s->alloc_pairs(std::max<int32_t>(2 * s->pair_cap, need + need / 4 + 16));
        grown = true;
    }
    if (status & ST_HITS) {
        s->alloc_hits(grown_hit_cap(s->d_hit_top.p, s->dpairs(par, false).hit_shards, s->hit_cap));
        grown = true;
    }
    if (status & ST_STAGE) {
        if (s->stage_cap >= (1 << 16)) return false;
        s->alloc_stage(s->stage_cap * 4);
        grown = true;
    }
    if (status & ST_TOPO) {
        // find what overflowed: the bump-allocated topology buffers, segments
        // beyond the walk scratch, or loops beyond lmax
        if (s->tw.grow_from_counters(s->st())) {
            s->drop_graphs();
            grown = true;
        }
        int32_t n = 0;
        HIPC(hipMemcpy(&n, s->ps[par].n.p, sizeof(int32_t), hipMemcpyDeviceToHost));
        std::vector<int32_t> ts(n), ns(n), nl(n);
        if (n) {
            HIPC(hipMemcpy(ts.data(), s->d_tstatus.p, n * sizeof(int32_t), hipMemcpyDeviceToHost));
            HIPC(hipMemcpy(ns.data(), s->d_nseg.p, n * sizeof(int32_t), hipMemcpyDeviceToHost));
        }
        int32_t smax = 0;
        bool loops = false;
        for (int p = 0; p < n; ++p) {
            if (ts[p] & TOPO_LOOP_OVERFLOW) loops = true;
            if (ts[p] & TOPO_E2_OVERFLOW) smax = std::max(smax, ns[p]);
        }
        if (loops) { s->alloc_lmax(s->lmax * 2); grown = true; }
        if (smax > s->seg_max) {
            s->seg_max = smax;
            s->alloc_topo_scratch();
            grown = true;
        }
    }
    if (status & ST_WALL) {
        s->wall_pts_cap *= 2;
        s->d_wscr.release();
        s->d_wscr.ensure((size_t)s->wall_threads * (size_t)(3 * s->wall_pts_cap + 4 * (s->wall_pts_cap + 8)));
        s->drop_graphs();
        grown = true;
    }
    if (status & ST_GRID) return false;
    return grown;
}

// Waits for the step's report: spins on rep->seq (set by the step's last
// kernel), asking the stream now and then whether it failed or drained.
void wait_report(zdx_sim* s) {
    volatile unsigned int* seq = &s->rep->seq;
    for (uint64_t spin = 1;; ++spin) {
        if (*seq) break;
        if ((spin & 255) == 0) {
            const hipError_t e = hipStreamQuery(s->st());
            if (e == hipSuccess) {
                if (*seq) break;
                throw ApiError{ZDX_HIP_ERROR, "step finished without its report"};
            }
            if (e != hipErrorNotReady) HIPC(e);
        }
        __builtin_ia32_pause();
    }
    std::atomic_thread_fence(std::memory_order_acquire);
}

void step_once(zdx_sim* s, bool batched) {
    hipStream_t st = s->st();
    s->ctx->sync_meshes();
    if (s->mesh_version != s->ctx->mesh_version) {
        s->drop_graphs();
        s->mesh_version = s->ctx->mesh_version;
    }
    for (int attempt = 0; attempt < 12; ++attempt) {
        const int par = s->parity;
        *(volatile unsigned int*)&s->rep->seq = 0u;
        std::atomic_thread_fence(std::memory_order_seq_cst);
        // ZDX_GRAPHS=0 runs the batched sequence without graph capture (same
        // kernels and launch shapes), for tools that cannot trace graphs
        static const bool use_graphs = [] {
            const char* e = std::getenv("ZDX_GRAPHS");
            return !(e && e[0] == '0');
        }();
        if (batched && !use_graphs) {
            enqueue_step(s, par, true);
        } else if (batched) {
            if (!s->graphs_valid) {
                s->acc_captures += 1;
                for (int k = 0; k < 2; ++k) {
                    HIPC(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
                    enqueue_step(s, k, true);
                    HIPC(hipStreamEndCapture(st, &s->graph[k]));
                    HIPC(hipGraphInstantiate(&s->gexec[k], s->graph[k], nullptr, nullptr, 0));
                }
                s->graphs_valid = true;
            }
            HIPC(hipGraphLaunch(s->gexec[par], st));
        } else {
            enqueue_step(s, par, false);
        }
        wait_report(s);
        const StepReport& r = *s->rep;
        const int32_t status = r.status;
        if (status == 0) {
            for (int k = 0; k < ZDX_N_COUNTERS; ++k) s->last_ctr[k] = (int64_t)r.ctr[k];
            const unsigned long long* kt = r.kt;
            // acc_ms: transform, broad phase, narrow, topology, walls (on its
            // own stream, overlapping narrow and topology), fold, narrow kernels
            const int a[7] = {1, 0, 2, 3, 9, 5, 8}, b[7] = {2, 1, 3, 4, 7, 6, 3};
            for (int k = 0; k < 7; ++k) {
                if (k == 4 && s->W == 0) continue;
                if (kt[b[k]] >= kt[a[k]]) s->acc_ms[k] += (double)(kt[b[k]] - kt[a[k]]) / s->clock_khz;
            }
            for (int k = 0; k < ZDX_N_COUNTERS; ++k)
                s->acc_ctr[k] = k == ZDX_CTR_SEG_MAX ? std::max(s->acc_ctr[k], (double)s->last_ctr[k])
                                                     : s->acc_ctr[k] + (double)s->last_ctr[k];
            s->acc_steps += 1;
            s->parity = 1 - par;
            s->n_steps_done += 1;
            return;
        }
        // the step wrote only the other state buffer and this step's pair set:
        // grow and run it again from the same state
        s->acc_retries += 1;
        if (!grow_for(s, status, par))
            throw ApiError{ZDX_CAPACITY, "simulation step capacity exceeded (status " + std::to_string(status) + ")"};
    }
    throw ApiError{ZDX_CAPACITY, "simulation step did not fit after repeated growth"};
}

}  // namespace

namespace {

// The per-pair call on the buffers of A: grown on demand (hipMalloc) for
// zdx_contact_pairs, carved from the context's no-allocation arena for
// zdx_contact_pairs_noalloc (lim then bounds the per-workgroup lists and the
// topology walk slices, which are sized by workgroup count).
void contact_pairs_impl(zdx_ctx* c, zdx_ctx::Api& A, const zdx_pair_in* pairs, int32_t n,
                        const zdx_contact_opts* opts, zdx_pair_out* out, const CallLimits& lim) {
    {
        HIPC(hipSetDevice(c->device));
        for (int p = 0; p < n; ++p)
            if (pairs[p].mesh_a < 0 || pairs[p].mesh_a >= (int)c->meshes.size() || pairs[p].mesh_b < 0 ||
                pairs[p].mesh_b >= (int)c->meshes.size())
                throw ApiError{ZDX_BAD_ARG, "pair references an unregistered mesh"};
        if (lim.may_allocate) c->sync_meshes();
        else if (c->mesh_dirty)
            throw ApiError{ZDX_BAD_ARG, "meshes registered since the last allocating call are not on the device yet "
                                        "(call zdx_contact_pairs once after registering)"};
        hipStream_t st = c->stream;
        const int nb = 2 * n;
        int64_t av = 0;
        for (int p = 0; p < n; ++p)
            av += (int64_t)(c->meshes[pairs[p].mesh_a].h.verts.size() / 3) +
                  (int64_t)(c->meshes[pairs[p].mesh_b].h.verts.size() / 3);
        if (av >= ((int64_t)1 << 31))   // the narrow phase's hit records index world vertices with int32
            throw ApiError{ZDX_CAPACITY, "more than 2^31 world vertices in one call"};
        if (n == 0) {
            out->loop_offset[0] = 0;
            if (out->seg_offset) out->seg_offset[0] = 0;
            out->n_loops_total = out->n_vtx_total = out->n_seg_total = 0;
            return;
        }
        // the pair records go to the device as they are; bodies, poses and
        // world-vertex offsets are expanded there
        A.pairs_in.ensure((size_t)n);
        HIPC(hipMemcpyAsync(A.pairs_in.p, pairs, (size_t)n * sizeof(zdx_pair_in), hipMemcpyHostToDevice, st));
        A.mesh.ensure(nb); A.state.ensure(16 * (size_t)nb); A.wv_off.ensure((size_t)nb + 1); A.wt_off.ensure((size_t)nb + 1);
        DMeshes M = c->dmeshes();
        launch_api_bodies(A.pairs_in.p, n, M, A.mesh.p, A.state.p, A.wv_off.p, A.wt_off.p, st);
        A.wv.ensure(3 * (size_t)av); A.aabb.ensure(6 * (size_t)nb); A.nondeg.ensure(nb);
        DBodies B;
        B.n = nb; B.mesh = A.mesh.p; B.state = A.state.p; B.wv_off = A.wv_off.p; B.wt_off = A.wt_off.p;
        B.wv = A.wv.p; B.nondeg = A.nondeg.p; B.aabb = A.aabb.p;
        HIPC(hipEventRecord(c->ev_api[0], st));
        launch_transform(B, M, st);
        HIPC(hipEventRecord(c->ev_api[1], st));
        A.pi.ensure(n); A.pj.ensure(n); A.npairs.ensure(1); A.status.ensure(1); A.order.ensure(n); A.items.ensure(n);
        A.status2.ensure(1);
        launch_iota_pairs(A.pi.p, A.pj.p, n, st);
        launch_fill_i32(A.npairs.p, n, 1, st);
        A.seg_off.ensure(n); A.tri_checks.ensure(n); A.tri_tests.ensure(n); A.nseg.ensure(n); A.tri_hits.ensure(n); A.tri_scanned.ensure(n);
        A.queue.ensure(NARROW_COUNTERS); A.heavy.ensure(n); A.bcnt.ensure(64); A.bcur.ensure(64);
        A.hit_off.ensure(n); A.hit_top.ensure(HIT_SHARDS * HIT_TOP_STRIDE);
        A.ncontrib.ensure(n); A.comps.ensure(n); A.loops_raw.ensure(n); A.tstatus.ensure(n);
        A.topo_list.ensure(n); A.topo_n.ensure(1);
        const bool batched = opts->mode == ZDX_MODE_BATCHED;
        int groups = std::min<int>(persistent_groups(c, 8), std::max(1, n));
        groups = (int)std::max<int64_t>(1, std::min<int64_t>(groups, lim.list_bytes / (80 * (int64_t)c->max_tris)));
        A.list_i.ensure((size_t)groups * 8 * c->max_tris);
        A.list_f.ensure((size_t)groups * 12 * c->max_tris);
        const bool want_vtx = out->loop_vtx != nullptr;
        for (int attempt = 0; attempt < 10; ++attempt) {
            c->api_attempts = attempt + 1;
            A.arena.ensure((size_t)A.hit_cap * 6);
            A.hits.ensure(6 * (size_t)A.hit_cap);
            A.hit_flag.ensure((size_t)A.hit_cap);
            const bool sens = opts->sensitivities != 0;
            DPairs P{};
            P.cap = n; P.n = A.npairs.p; P.pi = A.pi.p; P.pj = A.pj.p; P.order = batched ? A.order.p : nullptr;
            P.items = batched ? A.items.p : nullptr;
            P.seg_off = A.seg_off.p; P.nseg = A.nseg.p; P.tri_checks = A.tri_checks.p; P.tri_tests = A.tri_tests.p;
            P.tri_hits = A.tri_hits.p; P.tri_scanned = A.tri_scanned.p; P.arena = A.arena.p; P.arena_cap = A.hit_cap;
            P.status = A.status.p;
            P.topo_list = A.topo_list.p; P.topo_n = A.topo_n.p;
            P.hits = A.hits.p; P.hit_cap = A.hit_cap; P.hit_top = A.hit_top.p; P.hit_off = A.hit_off.p;
            P.hit_shards = (int32_t)std::min<int64_t>(std::min<int64_t>(HIT_SHARDS, n), A.hit_cap);
            P.hit_flag = A.hit_flag.p;
            P.heavy = A.heavy.p;
            if (sens) {
                A.hit_tri.ensure(2 * (size_t)A.hit_cap);
                P.hit_tri = A.hit_tri.p;
                A.prov.ensure((size_t)A.hit_cap * 4);
                P.prov = A.prov.p;
            }
            launch_fill_i32(A.status.p, 0, 1, st);
            launch_fill_i32(A.topo_n.p, 0, 1, st);
            HIPC(hipEventRecord(c->ev_api[2], st));
            if (batched) launch_lpt_order(P, M, B, A.order.p, A.bcnt.p, A.bcur.p, n, st);
            NarrowCfg nc{256, 4, groups, batched ? 1 : 0, std::min(persistent_groups(c, 4), (n + 3) / 4)};
            launch_narrow(B, M, P, nc, A.list_i.p, A.list_f.p, A.stage_cap, c->max_tris, A.queue.p, st);
            HIPC(hipEventRecord(c->ev_api[3], st));
            int32_t status = 0;
            HIPC(hipMemcpyAsync(&status, A.status.p, 4, hipMemcpyDeviceToHost, st));
            HIPC(hipStreamSynchronize(st));
            if (status & (ST_STAGE | ST_HITS)) {
                if (status & ST_STAGE) A.stage_cap *= 4;
                if (status & ST_HITS) A.hit_cap = grown_hit_cap(A.hit_top.p, P.hit_shards, A.hit_cap);
                continue;
            }
            // topology with the minimal-energy force
            TopoBufs T{};
            T.lmax = A.lmax; T.ncontrib = A.ncontrib.p; T.comps = A.comps.p; T.loops_raw = A.loops_raw.p;
            T.tstatus = A.tstatus.p;
            T.nthreads = std::min<int64_t>(lim.walk_slices, (int64_t)n);
            T.pass_groups = std::min<int64_t>(2048, (int64_t)n);
            T.seg_max = A.stage_cap;
            T.lds_bytes = 32 * 1024;
            T.vmax = 4 * A.stage_cap + 16;
            T.k_stiff = opts->k;
            T.tol_override = opts->tol > 0.0 ? opts->tol : 0.0;
            A.tw.ensure_arena(A.hit_cap);
            A.tw.ensure_items(n);
            A.tw.ensure_slices(T.nthreads, T.seg_max);
            A.tw.fill(T);
            A.rec.ensure((size_t)n * T.lmax * 28);
            if (want_vtx) A.vtx.ensure((size_t)n * T.vmax * 3);
            if (sens) {
                A.ej.ensure((size_t)A.hit_cap * 72);
                A.vj.ensure((size_t)A.hit_cap * 72);
                A.ljac.ensure((size_t)n * T.lmax * 72);
                T.ej = A.ej.p; T.vj = A.vj.p; T.ljac = A.ljac.p;
            }
            T.rec = A.rec.p; T.vtx = want_vtx ? A.vtx.p : nullptr;
            HIPC(hipEventRecord(c->ev_api[4], st));
            launch_topo_zero(P, T, nullptr, n, st);
            launch_topo_api(P, M, B, T, opts->split_contacts, st);
            HIPC(hipEventRecord(c->ev_api[5], st));
            HIPC(hipMemcpyAsync(&status, A.status.p, 4, hipMemcpyDeviceToHost, st));
            HIPC(hipStreamSynchronize(st));
            for (int k = 0; k < 3; ++k) {
                float ms = 0.0f;
                HIPC(hipEventElapsedTime(&ms, c->ev_api[2 * k], c->ev_api[2 * k + 1]));
                c->api_ms[k] = ms;
            }
            if (status & ST_TOPO) {
                int32_t ts = 0;
                launch_or_i32(A.tstatus.p, n, A.status2.p, st);
                HIPC(hipMemcpyAsync(&ts, A.status2.p, 4, hipMemcpyDeviceToHost, st));
                HIPC(hipStreamSynchronize(st));
                const bool lo = (ts & TOPO_LOOP_OVERFLOW) != 0, vo = (ts & TOPO_VTX_OVERFLOW) != 0;
                const bool grown = A.tw.grow_from_counters(st);
                if (lo) A.lmax *= 2;
                if (vo) A.stage_cap *= 2;
                if (!lo && !vo && !grown) throw ApiError{ZDX_CAPACITY, "contact topology exceeded its edge capacity"};
                continue;
            }
            // pack on the device, then one copy per output array
            const bool want_jac = sens && out->loop_jac;
            const bool want_seg = out->segments != nullptr;
            ApiPack K{};
            K.n = n; K.lmax = T.lmax; K.vmax = T.vmax;
            K.nloops = A.ncontrib.p; K.nseg = A.nseg.p; K.seg_off = A.seg_off.p;
            K.rec = A.rec.p; K.ljac = want_jac ? A.ljac.p : nullptr; K.vtx = want_vtx ? A.vtx.p : nullptr;
            K.arena = A.arena.p;
            A.pk_loop_off.ensure((size_t)n + 1); A.pk_seg_off.ensure((size_t)n + 1);
            A.pk_pvtx.ensure((size_t)n + 1); A.pk_tot.ensure(3);
            K.loop_off = A.pk_loop_off.p; K.seg_out_off = A.pk_seg_off.p; K.pvtx_off = A.pk_pvtx.p;
            K.totals = A.pk_tot.p;
            K.loop_cap = std::max<int64_t>(0, std::min<int64_t>(out->loop_cap, (int64_t)n * T.lmax));
            K.vtx_cap = want_vtx ? std::max<int64_t>(0, std::min<int64_t>(out->vtx_cap, (int64_t)n * T.vmax)) : 0;
            K.seg_cap = want_seg ? std::max<int64_t>(0, std::min<int64_t>(out->seg_cap, A.hit_cap)) : 0;
            if (out->loop_rec) { A.pk_rec.ensure((size_t)K.loop_cap * 28 + 1); K.rec_out = A.pk_rec.p; }
            if (want_jac) { A.pk_ljac.ensure((size_t)K.loop_cap * 72 + 1); K.ljac_out = A.pk_ljac.p; }
            if (want_vtx) {
                A.pk_vtx.ensure((size_t)K.vtx_cap * 3 + 1); K.vtx_out = A.pk_vtx.p;
                A.pk_vtx_off.ensure((size_t)K.loop_cap + 1); K.vtx_off_out = A.pk_vtx_off.p;
            }
            if (want_seg) { A.pk_seg.ensure((size_t)K.seg_cap * 6 + 1); K.seg_out = A.pk_seg.p; }
            launch_api_pack(K, st);
            int64_t tot[3] = {0, 0, 0};
            HIPC(hipMemcpyAsync(tot, A.pk_tot.p, sizeof(tot), hipMemcpyDeviceToHost, st));
            HIPC(hipStreamSynchronize(st));
            out->n_loops_total = tot[0];
            out->n_vtx_total = tot[1];
            out->n_seg_total = tot[2];
            if (tot[0] > out->loop_cap || (want_seg && tot[2] > out->seg_cap) || (want_vtx && tot[1] > out->vtx_cap))
                throw ApiError{ZDX_CAPACITY, "zdx_pair_out capacity too small (see n_*_total)"};
            if (tot[0] >= ((int64_t)1 << 31) || tot[1] >= ((int64_t)1 << 31) || tot[2] >= ((int64_t)1 << 31))
                throw ApiError{ZDX_CAPACITY, "more than 2^31 loops, loop vertices or segments in one call"};
            auto d2h = [&](void* dst, const void* src, size_t bytes) {
                if (bytes) HIPC(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st));
            };
            d2h(out->loop_offset, A.pk_loop_off.p, 4 * ((size_t)n + 1));
            if (out->tri_checks) d2h(out->tri_checks, A.tri_checks.p, 8 * (size_t)n);
            if (out->n_segments) d2h(out->n_segments, A.nseg.p, 4 * (size_t)n);
            if (out->n_comps) d2h(out->n_comps, A.comps.p, 4 * (size_t)n);
            if (out->n_loops_raw) d2h(out->n_loops_raw, A.loops_raw.p, 4 * (size_t)n);
            if (out->loop_rec) d2h(out->loop_rec, A.pk_rec.p, 8 * 28 * (size_t)tot[0]);
            if (want_jac) d2h(out->loop_jac, A.pk_ljac.p, 8 * 72 * (size_t)tot[0]);
            if (want_vtx) {
                d2h(out->vtx_offset, A.pk_vtx_off.p, 4 * ((size_t)tot[0] + 1));
                d2h(out->loop_vtx, A.pk_vtx.p, 8 * 3 * (size_t)tot[1]);
            }
            if (want_seg) {
                d2h(out->seg_offset, A.pk_seg_off.p, 4 * ((size_t)n + 1));
                d2h(out->segments, A.pk_seg.p, 8 * 6 * (size_t)tot[2]);
            }
            HIPC(hipStreamSynchronize(st));
            return;
        }
        throw ApiError{ZDX_CAPACITY, "per-pair buffers did not fit after repeated growth"};
    }
}

}  // namespace

// ======================================================================= C ABI
extern "C" {

zdx_status zdx_create(zdx_ctx** out, int32_t device) {
    if (!out) return ZDX_BAD_ARG;
    *out = nullptr;
    auto* c = new zdx_ctx();
    zdx_status st = guard(c, [&] {
        int n = 0;
        HIPC(hipGetDeviceCount(&n));
        if (device < 0 || device >= n) throw ApiError{ZDX_BAD_ARG, "no such HIP device"};
        HIPC(hipSetDevice(device));
        c->device = device;
        hipDeviceProp_t prop;
        HIPC(hipGetDeviceProperties(&prop, device));
        c->n_cu = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
        HIPC(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        HIPC(hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking));
        HIPC(hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
        HIPC(hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming));
        HIPC(hipStreamCreateWithFlags(&c->stream3, hipStreamNonBlocking));
        HIPC(hipEventCreateWithFlags(&c->ev_fork3, hipEventDisableTiming));
        HIPC(hipEventCreateWithFlags(&c->ev_join3, hipEventDisableTiming));
        HIPC(hipEventCreateWithFlags(&c->ev_cc3, hipEventDisableTiming));
        for (auto& e : c->ev_api) HIPC(hipEventCreate(&e));
        // this device's copy of the no-allocation scratch
        void* sp = nullptr;
        HIPC(hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch)));
        c->scratch_dev = sp;
        c->na_arena.base = (char*)sp;
        c->na_arena.cap = NOALLOC_SCRATCH_BYTES;
        if (tiny_start_capacity()) {
            auto& A = c->api;
            A.stage_cap = 4; A.hit_cap = 64; A.lmax = 1;
            A.tw.adj_cap = 64; A.tw.snap_cap = 64; A.tw.tj_cap = 64; A.tw.units_cap = 4;
        }
    });
    if (st != ZDX_OK) {
        delete c;
        return st;
    }
    *out = c;
    return ZDX_OK;
}

void zdx_destroy(zdx_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->stream2) (void)hipStreamSynchronize(c->stream2);
    if (c->stream3) (void)hipStreamSynchronize(c->stream3);
    for (auto& e : c->ev_api)
        if (e) (void)hipEventDestroy(e);
    if (c->ev_fork3) (void)hipEventDestroy(c->ev_fork3);
    if (c->ev_join3) (void)hipEventDestroy(c->ev_join3);
    if (c->ev_cc3) (void)hipEventDestroy(c->ev_cc3);
    if (c->stream3) (void)hipStreamDestroy(c->stream3);
    if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
    if (c->ev_join) (void)hipEventDestroy(c->ev_join);
    if (c->stream2) (void)hipStreamDestroy(c->stream2);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

const char* zdx_last_error(const zdx_ctx* c) { return c ? c->err.c_str() : "null context"; }

zdx_status zdx_load_stl(const char* path, double* tris, int32_t cap, int32_t* T) {
    if (!path || !T) return ZDX_BAD_ARG;
    std::vector<double> t;
    if (!load_stl(path, t)) return ZDX_IO;
    *T = (int32_t)(t.size() / 9);
    if (tris) {
        if (*T > cap) return ZDX_CAPACITY;
        std::memcpy(tris, t.data(), t.size() * sizeof(double));
    }
    return ZDX_OK;
}

zdx_status zdx_mesh_register(zdx_ctx* c, const double* tris, int32_t T, int32_t center, int32_t* mesh_id,
                             double* props) {
    if (!c || !tris || T <= 0 || !mesh_id) return ZDX_BAD_ARG;
    return guard(c, [&] {
        std::vector<double> t(tris, tris + 9 * (size_t)T);
        HostMesh h = build_mesh(t, center != 0);
        if (props) {
            props[0] = h.volume;
            std::memcpy(props + 1, h.center, 3 * sizeof(double));
            props[4] = h.radius; props[5] = h.mean_edge; props[6] = h.bbox_diag;
            std::memcpy(props + 7, h.inertia, 9 * sizeof(double));
        }
        *mesh_id = c->add_mesh(std::move(h));
    });
}

zdx_status zdx_contact_pairs(zdx_ctx* c, const zdx_pair_in* pairs, int32_t n, const zdx_contact_opts* opts,
                             zdx_pair_out* out) {
    if (!c || (!pairs && n > 0) || n < 0 || !opts || !out || !out->loop_offset) return ZDX_BAD_ARG;
    return guard(c, [&] { contact_pairs_impl(c, c->api, pairs, n, opts, out, CallLimits{}); });
}

zdx_status zdx_contact_pairs_noalloc(zdx_ctx* c, const zdx_pair_in* pairs, int32_t n, const zdx_contact_opts* opts,
                                     zdx_pair_out* out) {
    if (!c || (!pairs && n > 0) || n < 0 || !opts || !out || !out->loop_offset) return ZDX_BAD_ARG;
    return guard(c, [&] {
        if (!c->scratch_dev) throw ApiError{ZDX_INTERNAL, "no-allocation scratch address not set up"};
        // every call starts from the front of this device's scratch array
        c->na_arena.reset();
        c->api_na.bind(&c->na_arena);
        CallLimits lim;
        lim.may_allocate = false;
        lim.list_bytes = NOALLOC_LIST_BYTES;
        lim.walk_slices = NOALLOC_WALK_SLICES;
        contact_pairs_impl(c, c->api_na, pairs, n, opts, out, lim);
    });
}

zdx_status zdx_wall_contacts(zdx_ctx* c, const zdx_body_pose* bodies, int32_t n, const double* wall_tris,
                             const double* wall_normals, int32_t n_tris, int32_t jac, double* out) {
    if (!c || n < 0 || n_tris < 0 || ((n > 0 && n_tris > 0) && (!bodies || !wall_tris || !wall_normals || !out)))
        return ZDX_BAD_ARG;
    return guard(c, [&] {
        HIPC(hipSetDevice(c->device));
        if (n == 0 || n_tris == 0) return;
        for (int b = 0; b < n; ++b)
            if (bodies[b].mesh < 0 || bodies[b].mesh >= (int)c->meshes.size())
                throw ApiError{ZDX_BAD_ARG, "body references an unregistered mesh"};
        c->sync_meshes();
        hipStream_t st = c->stream;
        auto& Q = c->wq;
        std::vector<int32_t> mesh(n);
        std::vector<double> state(16 * (size_t)n, 0.0);
        std::vector<int64_t> wvo(n + 1), wto(n + 1);
        int64_t av = 0, at = 0;
        int32_t tmax = 1;
        for (int b = 0; b < n; ++b) {
            mesh[b] = bodies[b].mesh;
            for (int k = 0; k < 3; ++k) state[16 * (size_t)b + k] = bodies[b].pos[k];
            for (int k = 0; k < 4; ++k) state[16 * (size_t)b + 3 + k] = bodies[b].quat[k];
            wvo[b] = av; wto[b] = at;
            av += (int64_t)(c->meshes[mesh[b]].h.verts.size() / 3);
            const int32_t T = (int32_t)(c->meshes[mesh[b]].h.tri_idx.size() / 3);
            at += T;
            tmax = std::max(tmax, T);
        }
        wvo[n] = av; wto[n] = at;
        Q.mesh.upload(mesh.data(), n, st);
        Q.state.upload(state.data(), state.size(), st);
        Q.wv_off.upload(wvo.data(), n + 1, st);
        Q.wt_off.upload(wto.data(), n + 1, st);
        Q.wv.ensure(3 * (size_t)av); Q.aabb.ensure(6 * (size_t)n); Q.nondeg.ensure(n);
        Q.tris.upload(wall_tris, 9 * (size_t)n_tris, st);
        Q.normals.upload(wall_normals, 3 * (size_t)n_tris, st);
        DBodies B;
        B.n = n; B.mesh = Q.mesh.p; B.state = Q.state.p; B.wv_off = Q.wv_off.p; B.wt_off = Q.wt_off.p;
        B.wv = Q.wv.p; B.nondeg = Q.nondeg.p; B.aabb = Q.aabb.p;
        DMeshes M = c->dmeshes();
        launch_transform(B, M, st);
        // section points: at most one per edge slot of the body mesh
        const int32_t cap = 3 * tmax;
        const int64_t nq = (int64_t)n * n_tris;
        Q.scr.ensure((size_t)nq * (size_t)(3 * (int64_t)cap + 32 * (int64_t)(cap + 8)));
        Q.isrc.ensure((size_t)nq * (size_t)(cap + 1));
        Q.out.ensure(32 * (size_t)nq);
        Q.status.ensure(1);
        launch_fill_i32(Q.status.p, 0, 1, st);
        launch_wall_query(B, M, n, Q.tris.p, Q.normals.p, n_tris, jac ? 1 : 0, cap, Q.scr.p, Q.isrc.p, Q.out.p,
                          Q.status.p, st);
        int32_t status = 0;
        HIPC(hipMemcpyAsync(&status, Q.status.p, 4, hipMemcpyDeviceToHost, st));
        HIPC(hipMemcpyAsync(out, Q.out.p, 32 * 8 * (size_t)nq, hipMemcpyDeviceToHost, st));
        HIPC(hipStreamSynchronize(st));
        if (status & 2) throw ApiError{ZDX_INTERNAL, "wall contact derivative disagreed with the contact"};
        if (status) throw ApiError{ZDX_INTERNAL, "wall section exceeded its point capacity"};
    });
}

zdx_status zdx_contact_times(zdx_ctx* c, double* out, int32_t n) {
    if (!c || (!out && n > 0) || n < 0) return ZDX_BAD_ARG;
    const double v[4] = {c->api_ms[0], c->api_ms[1], c->api_ms[2], (double)c->api_attempts};
    for (int k = 0; k < n && k < 4; ++k) out[k] = v[k];
    return ZDX_OK;
}

zdx_status zdx_sim_create(zdx_ctx* c, const zdx_sim_desc* d, zdx_sim** out) {
    if (!c || !d || !out || d->n_particles <= 0 || !d->mesh_id || !d->init || !d->material) return ZDX_BAD_ARG;
    *out = nullptr;
    auto* s = new zdx_sim();
    zdx_status st = guard(c, [&] {
        HIPC(hipSetDevice(c->device));
        s->ctx = c;
        s->N = d->n_particles;
        for (int i = 0; i < s->N; ++i)
            if (d->mesh_id[i] < 0 || d->mesh_id[i] >= (int)c->meshes.size())
                throw ApiError{ZDX_BAD_ARG, "particle references an unregistered mesh"};
        s->sp.g[0] = d->gravity[0]; s->sp.g[1] = d->gravity[1]; s->sp.g[2] = d->gravity[2];
        s->sp.dt = d->dt;
        s->sp.split = d->split_contacts;
        s->pc.resize(s->N);
        s->mesh_of.resize(s->N);
        std::vector<double> state(16 * (size_t)s->N);
        for (int i = 0; i < s->N; ++i) {
            const double* m = d->material + 6 * (size_t)i;
            s->pc[i] = particle_const(c->meshes[d->mesh_id[i]].h, d->mesh_id[i], m[0], m[1], m[2], m[3], m[4]);
            s->mesh_of[i] = d->mesh_id[i];
            const double* in = d->init + 13 * (size_t)i;
            double* x = &state[16 * (size_t)i];
            std::memcpy(x, in, 13 * sizeof(double));
            initial_L(s->pc[i], in + 3, in + 10, x + 13);
        }
        s->inits.resize(s->N);
        for (int i = 0; i < s->N; ++i) {
            const double* m = d->material + 6 * (size_t)i;
            ParticleInit& pi = s->inits[i];
            pi.scale = 1.0; pi.density = m[0]; pi.young = m[1]; pi.poisson = m[2]; pi.mu = m[3]; pi.rest = m[4];
        }
        sim_setup_device(s);
        std::vector<WallHost> walls;
        for (int w = 0; w < d->n_walls; ++w) {
            int mid = d->wall_mesh_id[w];
            if (mid < 0 || mid >= (int)c->meshes.size()) throw ApiError{ZDX_BAD_ARG, "wall references an unregistered mesh"};
            walls.push_back(make_wall(c->meshes[mid].h, d->wall_pose + 7 * (size_t)w, d->wall_pose + 7 * (size_t)w + 3,
                                      d->wall_material[2 * w], d->wall_material[2 * w + 1]));
        }
        sim_setup_walls(s, walls);
        HIPC(hipMemcpy(s->d_state[0].p, state.data(), state.size() * 8, hipMemcpyHostToDevice));
    });
    if (st != ZDX_OK) {
        delete s;
        return st;
    }
    *out = s;
    return ZDX_OK;
}

zdx_status zdx_sim_load_config(zdx_ctx* c, const char* path, const char* base_dir, zdx_sim** out) {
    if (!c || !path || !out) return ZDX_BAD_ARG;
    *out = nullptr;
    SimConfig cfg;
    if (!parse_config(path, cfg)) {
        c->err = std::string("failed to parse config ") + path;
        return ZDX_IO;
    }
    std::vector<int32_t> mesh_ids;
    std::vector<double> init, mat;
    std::vector<int32_t> wmesh;
    std::vector<double> wpose, wmat;
    ConfigScene sc;
    std::string err;
    if (!build_scene(cfg, base_dir ? base_dir : "", sc, err)) {
        c->err = err;
        return err.rfind("failed to load", 0) == 0 ? ZDX_IO : ZDX_BAD_ARG;
    }
    zdx_status st = guard(c, [&] {
        std::vector<int> ctx_id(sc.meshes.size());
        for (size_t m = 0; m < sc.meshes.size(); ++m) ctx_id[m] = c->add_mesh(std::move(sc.meshes[m]));
        for (size_t i = 0; i < sc.inits.size(); ++i) {
            const ParticleInit& in = sc.inits[i];
            mesh_ids.push_back(ctx_id[sc.mesh_of[i]]);
            init.insert(init.end(), in.pos, in.pos + 3);
            init.insert(init.end(), in.quat, in.quat + 4);
            init.insert(init.end(), in.vel, in.vel + 3);
            init.insert(init.end(), in.omega, in.omega + 3);
            double m[6] = {in.density, in.young, in.poisson, in.mu, in.rest, 0.0};
            mat.insert(mat.end(), m, m + 6);
        }
        for (size_t w = 0; w < sc.walls.size(); ++w) {
            const WallInit& wi = sc.walls[w];
            wmesh.push_back(c->add_mesh(std::move(sc.wall_meshes[w])));
            wpose.insert(wpose.end(), wi.pos, wi.pos + 3);
            wpose.insert(wpose.end(), wi.quat, wi.quat + 4);
            wmat.push_back(wi.mu);
            wmat.push_back(wi.rest);
        }
    });
    if (st != ZDX_OK) return st;
    std::vector<ParticleInit>& inits = sc.inits;
    zdx_sim_desc d{};
    d.n_particles = (int32_t)mesh_ids.size();
    d.mesh_id = mesh_ids.data();
    d.init = init.data();
    d.material = mat.data();
    d.n_walls = (int32_t)wmesh.size();
    d.wall_mesh_id = wmesh.data();
    d.wall_pose = wpose.data();
    d.wall_material = wmat.data();
    std::memcpy(d.gravity, cfg.gravity, sizeof(d.gravity));
    d.dt = cfg.dt;
    d.split_contacts = cfg.split ? 1 : 0;
    st = zdx_sim_create(c, &d, out);
    if (st == ZDX_OK) {
        (*out)->cfg = cfg;
        (*out)->inits = inits;
        if (cfg.force_model != 0) st = zdx_sim_set_force_model(*out, cfg.force_model, cfg.k_minimal);
        if (st != ZDX_OK) {
            zdx_sim_destroy(*out);
            *out = nullptr;
        }
    }
    return st;
}

void zdx_sim_destroy(zdx_sim* s) {
    if (!s) return;
    (void)hipSetDevice(s->ctx->device);
    (void)hipStreamSynchronize(s->ctx->stream);
    delete s;
}

int32_t zdx_sim_num_particles(const zdx_sim* s) { return s ? s->N : 0; }

zdx_status zdx_sim_step(zdx_sim* s, int32_t n_steps, int32_t mode, int64_t* counters) {
    if (!s || n_steps < 0 || (mode != ZDX_MODE_PLAIN && mode != ZDX_MODE_BATCHED)) return ZDX_BAD_ARG;
    return guard(s->ctx, [&] {
        HIPC(hipSetDevice(s->ctx->device));
        for (int k = 0; k < n_steps; ++k) step_once(s, mode == ZDX_MODE_BATCHED);
        if (counters) std::memcpy(counters, s->last_ctr, sizeof(s->last_ctr));
    });
}

zdx_status zdx_sim_get_state(zdx_sim* s, double* state) {
    if (!s || !state) return ZDX_BAD_ARG;
    return guard(s->ctx, [&] {
        HIPC(hipStreamSynchronize(s->st()));
        HIPC(hipMemcpy(state, s->d_state[s->parity].p, 16 * (size_t)s->N * 8, hipMemcpyDeviceToHost));
    });
}

zdx_status zdx_sim_set_state(zdx_sim* s, const double* state) {
    if (!s || !state) return ZDX_BAD_ARG;
    return guard(s->ctx, [&] {
        HIPC(hipStreamSynchronize(s->st()));
        HIPC(hipMemcpy(s->d_state[s->parity].p, state, 16 * (size_t)s->N * 8, hipMemcpyHostToDevice));
    });
}

zdx_status zdx_sim_get_forces(zdx_sim* s, double* F, double* T, int32_t* cnt) {
    if (!s) return ZDX_BAD_ARG;
    return guard(s->ctx, [&] {
        HIPC(hipStreamSynchronize(s->st()));
        if (F) HIPC(hipMemcpy(F, s->d_F.p, 3 * (size_t)s->N * 8, hipMemcpyDeviceToHost));
        if (T) HIPC(hipMemcpy(T, s->d_T.p, 3 * (size_t)s->N * 8, hipMemcpyDeviceToHost));
        if (cnt) HIPC(hipMemcpy(cnt, s->d_counts.p, (size_t)s->N * 4, hipMemcpyDeviceToHost));
    });
}

zdx_status zdx_sim_get_pairs(zdx_sim* s, int32_t* pi, int32_t* pj, int64_t cap, int64_t* n) {
    if (!s || !n || cap < 0 || (cap > 0 && (!pi || !pj))) return ZDX_BAD_ARG;
    return guard(s->ctx, [&] {
        HIPC(hipStreamSynchronize(s->st()));
        *n = 0;
        if (s->n_steps_done == 0) return;
        const int last = 1 - s->parity;   // the pair set the last completed step wrote
        int32_t np = 0;
        HIPC(hipMemcpy(&np, s->ps[last].n.p, sizeof(int32_t), hipMemcpyDeviceToHost));
        *n = np;
        if (np > cap) throw ApiError{ZDX_CAPACITY, "pair buffers too small (see *n)"};
        if (np > 0) {
            HIPC(hipMemcpy(pi, s->ps[last].pi.p, (size_t)np * 4, hipMemcpyDeviceToHost));
            HIPC(hipMemcpy(pj, s->ps[last].pj.p, (size_t)np * 4, hipMemcpyDeviceToHost));
        }
    });
}

zdx_status zdx_sim_write_vtk(zdx_sim* s, const char* path) {
    if (!s || !path) return ZDX_BAD_ARG;
    return guard(s->ctx, [&] {
        const size_t n = (size_t)s->N;
        std::vector<double> state(16 * n), F(3 * n), T(3 * n), mass(n), radius(n);
        std::vector<int32_t> cnt(n);
        std::vector<const HostMesh*> mesh(n);
        HIPC(hipStreamSynchronize(s->st()));
        HIPC(hipMemcpy(state.data(), s->d_state[s->parity].p, state.size() * 8, hipMemcpyDeviceToHost));
        HIPC(hipMemcpy(F.data(), s->d_F.p, F.size() * 8, hipMemcpyDeviceToHost));
        HIPC(hipMemcpy(T.data(), s->d_T.p, T.size() * 8, hipMemcpyDeviceToHost));
        HIPC(hipMemcpy(cnt.data(), s->d_counts.p, cnt.size() * 4, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < n; ++i) {
            mass[i] = s->pc[i].mass;
            radius[i] = s->pc[i].radius;
            mesh[i] = &s->ctx->meshes[s->mesh_of[i]].h;
        }
        if (!write_vtk_particles(path, (int)n, mesh.data(), state.data(), mass.data(), radius.data(), F.data(),
                                 T.data(), cnt.data()))
            throw ApiError{ZDX_IO, std::string("failed to write VTK ") + path};
    });
}

zdx_status zdx_sim_write_state_txt(zdx_sim* s, const char* path) {
    if (!s || !path) return ZDX_BAD_ARG;
    return guard(s->ctx, [&] {
        std::vector<double> state(16 * (size_t)s->N);
        HIPC(hipStreamSynchronize(s->st()));
        HIPC(hipMemcpy(state.data(), s->d_state[s->parity].p, state.size() * 8, hipMemcpyDeviceToHost));
        if (!write_particle_state_txt(path, s->N, state.data(), s->inits.data()))
            throw ApiError{ZDX_IO, std::string("failed to write particle state ") + path};
    });
}

// Checkpoint file: magic, N, W, parity-independent payload: state (16 N),
// pair history (n, keys, ds, active) of the last step, wall history (3 N W ds,
// N W active).
zdx_status zdx_sim_checkpoint(zdx_sim* s, const char* path) {
    if (!s || !path) return ZDX_BAD_ARG;
    return guard(s->ctx, [&] {
        HIPC(hipStreamSynchronize(s->st()));
        const int prev = 1 - s->parity;  // set written by the last completed step
        std::vector<double> state(16 * (size_t)s->N);
        HIPC(hipMemcpy(state.data(), s->d_state[s->parity].p, state.size() * 8, hipMemcpyDeviceToHost));
        int32_t n = 0;
        HIPC(hipMemcpy(&n, s->ps[prev].n.p, 4, hipMemcpyDeviceToHost));
        std::vector<int64_t> key(n);
        std::vector<double> ds(3 * (size_t)n);
        std::vector<uint8_t> act(n);
        if (n) {
            HIPC(hipMemcpy(key.data(), s->ps[prev].key.p, n * 8, hipMemcpyDeviceToHost));
            HIPC(hipMemcpy(ds.data(), s->ps[prev].ds.p, ds.size() * 8, hipMemcpyDeviceToHost));
            HIPC(hipMemcpy(act.data(), s->ps[prev].active.p, n, hipMemcpyDeviceToHost));
        }
        const size_t npw = (size_t)s->N * s->W;
        std::vector<double> wds(3 * npw);
        std::vector<uint8_t> wact(npw);
        if (npw) {
            HIPC(hipMemcpy(wds.data(), s->ps[prev].wds.p, wds.size() * 8, hipMemcpyDeviceToHost));
            HIPC(hipMemcpy(wact.data(), s->ps[prev].wactive.p, npw, hipMemcpyDeviceToHost));
        }
        FILE* f = std::fopen(path, "wb");
        if (!f) throw ApiError{ZDX_IO, std::string("cannot write ") + path};
        const char magic[8] = {'Z', 'D', 'X', 'C', 'K', 'P', '1', 0};
        int64_t hdr[4] = {s->N, s->W, n, s->n_steps_done};
        bool ok = std::fwrite(magic, 1, 8, f) == 8 && std::fwrite(hdr, 8, 4, f) == 4 &&
                  std::fwrite(state.data(), 8, state.size(), f) == state.size() &&
                  std::fwrite(key.data(), 8, key.size(), f) == key.size() &&
                  std::fwrite(ds.data(), 8, ds.size(), f) == ds.size() &&
                  std::fwrite(act.data(), 1, act.size(), f) == act.size() &&
                  std::fwrite(wds.data(), 8, wds.size(), f) == wds.size() &&
                  std::fwrite(wact.data(), 1, wact.size(), f) == wact.size();
        std::fclose(f);
        if (!ok) throw ApiError{ZDX_IO, "checkpoint write failed"};
    });
}

zdx_status zdx_sim_restore(zdx_sim* s, const char* path) {
    if (!s || !path) return ZDX_BAD_ARG;
    return guard(s->ctx, [&] {
        HIPC(hipStreamSynchronize(s->st()));
        FILE* f = std::fopen(path, "rb");
        if (!f) throw ApiError{ZDX_IO, std::string("cannot read ") + path};
        char magic[8];
        int64_t hdr[4];
        bool ok = std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, "ZDXCKP1", 7) == 0 &&
                  std::fread(hdr, 8, 4, f) == 4 && hdr[0] == s->N && hdr[1] == s->W;
        if (!ok) { std::fclose(f); throw ApiError{ZDX_IO, "checkpoint does not match this simulation"}; }
        const int32_t n = (int32_t)hdr[2];
        std::vector<double> state(16 * (size_t)s->N);
        std::vector<int64_t> key(n);
        std::vector<double> ds(3 * (size_t)n);
        std::vector<uint8_t> act(n);
        const size_t npw = (size_t)s->N * s->W;
        std::vector<double> wds(3 * npw);
        std::vector<uint8_t> wact(npw);
        ok = std::fread(state.data(), 8, state.size(), f) == state.size() &&
             std::fread(key.data(), 8, key.size(), f) == key.size() &&
             std::fread(ds.data(), 8, ds.size(), f) == ds.size() &&
             std::fread(act.data(), 1, act.size(), f) == act.size() &&
             std::fread(wds.data(), 8, wds.size(), f) == wds.size() &&
             std::fread(wact.data(), 1, wact.size(), f) == wact.size();
        std::fclose(f);
        if (!ok) throw ApiError{ZDX_IO, "checkpoint truncated"};
        if (n > s->pair_cap) s->alloc_pairs(n + n / 4 + 16);
        const int prev = 1 - s->parity;
        HIPC(hipMemcpy(s->d_state[s->parity].p, state.data(), state.size() * 8, hipMemcpyHostToDevice));
        HIPC(hipMemcpy(s->ps[prev].n.p, &n, 4, hipMemcpyHostToDevice));
        if (n) {
            HIPC(hipMemcpy(s->ps[prev].key.p, key.data(), n * 8, hipMemcpyHostToDevice));
            HIPC(hipMemcpy(s->ps[prev].ds.p, ds.data(), ds.size() * 8, hipMemcpyHostToDevice));
            HIPC(hipMemcpy(s->ps[prev].active.p, act.data(), n, hipMemcpyHostToDevice));
        }
        if (npw) {
            HIPC(hipMemcpy(s->ps[prev].wds.p, wds.data(), wds.size() * 8, hipMemcpyHostToDevice));
            HIPC(hipMemcpy(s->ps[prev].wactive.p, wact.data(), npw, hipMemcpyHostToDevice));
        }
        s->n_steps_done = hdr[3];
    });
}

zdx_status zdx_sim_stage_times(zdx_sim* s, double* out, int32_t n) {
    if (!s || !out || n < 0) return ZDX_BAD_ARG;
    double v[10 + ZDX_N_COUNTERS];
    const double steps = (double)s->acc_steps;
    for (int k = 0; k < 6; ++k) v[k] = steps > 0 ? s->acc_ms[k] / steps : 0.0;
    v[6] = steps;
    for (int k = 0; k < ZDX_N_COUNTERS; ++k) v[7 + k] = s->acc_ctr[k];
    v[7 + ZDX_N_COUNTERS] = steps > 0 ? s->acc_ms[6] / steps : 0.0;
    v[8 + ZDX_N_COUNTERS] = (double)s->acc_retries;
    v[9 + ZDX_N_COUNTERS] = (double)s->acc_captures;
    for (int k = 0; k < n && k < 10 + ZDX_N_COUNTERS; ++k) out[k] = v[k];
    for (int k = 0; k < 7; ++k) s->acc_ms[k] = 0.0;
    for (int k = 0; k < ZDX_N_COUNTERS; ++k) s->acc_ctr[k] = 0.0;
    s->acc_steps = 0;
    s->acc_retries = 0;
    s->acc_captures = 0;
    return ZDX_OK;
}

zdx_status zdx_sim_set_force_model(zdx_sim* s, int32_t model, double k) {
    if (!s || (model != ZDX_FORCE_ZDEM && model != ZDX_FORCE_MINIMAL && model != ZDX_FORCE_ZDEM_DAMPED))
        return ZDX_BAD_ARG;
    if (model == ZDX_FORCE_MINIMAL && !(k > 0.0 && k < 1e300)) {
        s->ctx->err = "the minimal-energy force needs a finite stiffness k > 0";
        return ZDX_BAD_ARG;
    }
    s->sp.force_model = model;
    s->sp.k_minimal = model == ZDX_FORCE_MINIMAL ? k : 0.0;
    s->drop_graphs();   // the graphs hold the step parameters by value
    return ZDX_OK;
}

zdx_status zdx_sim_set_tiles(zdx_sim* s, int32_t shape, int32_t groups_per_cu) {
    if (!s) return ZDX_BAD_ARG;
    // the narrow-phase kernel shapes that are compiled (launch_narrow)
    static const int32_t shapes[] = {6404, 6408, 12804, 12808, 25601, 25602, 25604, 25608, 51202, 51204};
    if (shape > 0 && std::find(std::begin(shapes), std::end(shapes), shape) == std::end(shapes)) {
        s->ctx->err = "narrow-phase shape " + std::to_string(shape) +
                      " is not threads*100+tests per thread of a compiled kernel (6404 6408 12804 12808 25601 "
                      "25602 25604 25608 51202 51204)";
        return ZDX_BAD_ARG;
    }
    if (groups_per_cu < 0 || groups_per_cu > 16) {
        s->ctx->err = "workgroups per CU must be 1..16 (0 keeps the current value)";
        return ZDX_BAD_ARG;
    }
    if (shape > 0) {
        s->ncfg.block = shape / 100;
        s->ncfg.items = shape % 100;
    }
    if (groups_per_cu > 0) s->np_groups_per_cu = groups_per_cu;
    s->drop_graphs();
    return ZDX_OK;
}

// Autotuner: times the narrow phase alone on the current step's pair set for
// each (threads per workgroup, tests per thread, workgroups per CU of the
// workgroup kernel; most tests of a pair on the one-wave kernel) and keeps the
// fastest.  Every configuration produces the identical segment lists.
zdx_status zdx_sim_autotune(zdx_sim* s, int32_t* out_block, int32_t* out_groups) {
    if (!s) return ZDX_BAD_ARG;
    return guard(s->ctx, [&] {
        hipStream_t st = s->st();
        const int par = s->parity;
        DBodies B = s->bodies(par);
        DMeshes M = s->ctx->dmeshes();
        launch_fill_i32(s->d_status.p, 0, 1, st);
        launch_broadphase(s->bpbufs(par), st);
        HIPC(hipStreamSynchronize(st));
        if (read_status(s) & ST_PAIRS) {
            grow_for(s, ST_PAIRS, par);
            launch_fill_i32(s->d_status.p, 0, 1, st);
            launch_broadphase(s->bpbufs(par), st);
        }
        launch_need_list(s->N, s->d_state[par].p, s->d_radius.p, s->d_cnt_all.p, s->wallbufs(), s->d_need_list.p,
                         s->d_n_need.p, s->d_wall_list.p, s->d_n_wall.p, false, st);
        launch_transform(B, M, st, s->d_need_list.p, s->d_n_need.p, s->transform_groups());
        DPairs P = s->dpairs(par, true);
        P.topo_big = P.topo_small = P.n_big = P.n_small = nullptr;   // timing the narrow phase only
        launch_lpt_order(P, M, B, s->d_order.p, s->d_bcnt.p, s->d_bcur.p, s->pair_cap, st);
        const int cand[][2] = {{64, 4}, {64, 8}, {128, 4}, {128, 8}, {256, 1}, {256, 2}, {256, 4}, {256, 8},
                               {512, 2}, {512, 4}};
        const int gpc[] = {4, 8, 16};
        hipEvent_t e0, e1;
        HIPC(hipEventCreate(&e0));
        HIPC(hipEventCreate(&e1));
        float best = 1e30f;
        NarrowCfg bestc = s->ncfg;
        int bestg = s->np_groups_per_cu;
        // and the most float-box tests of a pair on the one-wave kernel
        const int wtests[] = {4096, 16384};
        for (auto& cb : cand) {
            for (int g : gpc) for (int wt : wtests) {
                NarrowCfg nc{cb[0], cb[1], persistent_groups(s->ctx, g), 1, persistent_groups(s->ctx, 4), wt};
                // one untimed launch, then the median of four timed ones
                float t4[4] = {0, 0, 0, 0};
                for (int rep = 0; rep < 5; ++rep) {
                    launch_fill_i32(s->d_topo_n.p, 0, 1, st);
                    HIPC(hipEventRecord(e0, st));
                    launch_narrow(B, M, P, nc, s->d_list_i.p, s->d_list_f.p, s->stage_cap, s->tmax, s->d_queue.p,
                                  st);
                    HIPC(hipEventRecord(e1, st));
                    HIPC(hipEventSynchronize(e1));
                    float ms = 0;
                    HIPC(hipEventElapsedTime(&ms, e0, e1));
                    if (rep > 0) t4[rep - 1] = ms;
                }
                std::sort(t4, t4 + 4);
                const float tot = 0.5f * (t4[1] + t4[2]);
                if (tot < best) { best = tot; bestc = nc; bestg = g; }
            }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        s->ncfg.block = bestc.block;
        s->ncfg.items = bestc.items;
        s->ncfg.wave_tests = bestc.wave_tests;
        s->np_groups_per_cu = bestg;
        s->drop_graphs();
        if (out_block) *out_block = bestc.block * 100 + bestc.items;
        if (out_groups) *out_groups = bestg;
    });
}

}  // extern "C"
