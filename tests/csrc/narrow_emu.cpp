// narrow_emu.cpp — the narrow kernels of zedm-x_amd/csrc/zdx_narrow.h run on
// the CPU under the SIMT emulation of simt_emu.h: the one-wave-per-pair
// kernel k_narrow_wave followed by the workgroup kernel k_narrow over the
// pairs it leaves, as the library launches them, and k_narrow alone over every
// pair (each of its paths), for every compiled workgroup shape, in the time
// step's form (work items in a shuffled order from the work queue) and the
// per-pair call's (pairs in index order, striding the grid), with and without
// (ta, tb) provenance.
//
// Each run is checked against a plain restatement of what the kernel must
// produce: per pair, A's triangles whose outward-rounded float box meets B's
// float AABB, B's likewise against A's, and the pairs (ta, tb) of the two
// lists, in (ta, tb) order, whose float boxes overlap -- with the rounding
// done here by the C library under FE_DOWNWARD / FE_UPWARD, independent of
// both the kernel's outcode cull and the emulation's conversions.  Hit
// records (six world-vertex indices), provenance, per-pair counters, the
// hit arena's packing and the status word must all match.
//
//   narrow_emu INPUT     (binary written by tests/test_narrow_emu.py)
#include "simt_emu.h"

#include <algorithm>
#include <cfenv>
#include <numeric>
#include <random>

#include "../../zedm-x_amd/csrc/zdx_geom.h"
#include "../../zedm-x_amd/csrc/zdx_internal.h"

namespace zdx {
namespace {
constexpr int WAVE = 64;
#include "../../zedm-x_amd/csrc/zdx_wave.h"
#include "../../zedm-x_amd/csrc/zdx_narrow.h"
}  // namespace
}  // namespace zdx

using namespace zdx;

struct Input {
    std::vector<int32_t> nverts, ntris, tri_off, tri_idx;
    std::vector<int32_t> mesh, nondeg;
    std::vector<int64_t> wv_off;
    std::vector<double> wv, aabb;
    std::vector<int32_t> pi, pj;
};

static bool read_input(const char* path, Input& in) {
    FILE* f = std::fopen(path, "rb");
    if (!f) return false;
    auto rd = [&](void* p, size_t n) { return std::fread(p, 1, n, f) == n; };
    int32_t nm = 0;
    if (!rd(&nm, 4)) return false;
    for (int m = 0; m < nm; ++m) {
        int32_t nv, nt;
        if (!rd(&nv, 4) || !rd(&nt, 4)) return false;
        in.nverts.push_back(nv);
        in.ntris.push_back(nt);
        in.tri_off.push_back((int32_t)(in.tri_idx.size() / 3));
        const size_t o = in.tri_idx.size();
        in.tri_idx.resize(o + 3 * (size_t)nt);
        if (!rd(in.tri_idx.data() + o, 12 * (size_t)nt)) return false;
    }
    int32_t nb = 0;
    if (!rd(&nb, 4)) return false;
    for (int b = 0; b < nb; ++b) {
        int32_t m, nd;
        if (!rd(&m, 4) || !rd(&nd, 4)) return false;
        in.mesh.push_back(m);
        in.nondeg.push_back(nd);
        in.wv_off.push_back((int64_t)(in.wv.size() / 3));
        const size_t o = in.wv.size();
        const int nv = in.nverts[m];
        in.wv.resize(o + 3 * (size_t)nv);
        if (!rd(in.wv.data() + o, 24 * (size_t)nv)) return false;
        double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int v = 0; v < nv; ++v)
            for (int k = 0; k < 3; ++k) {
                lo[k] = std::min(lo[k], in.wv[o + 3 * v + k]);
                hi[k] = std::max(hi[k], in.wv[o + 3 * v + k]);
            }
        for (int k = 0; k < 3; ++k) in.aabb.push_back(lo[k]);
        for (int k = 0; k < 3; ++k) in.aabb.push_back(hi[k]);
    }
    int32_t np = 0;
    if (!rd(&np, 4)) return false;
    in.pi.resize(np);
    in.pj.resize(np);
    for (int p = 0; p < np; ++p)
        if (!rd(&in.pi[p], 4) || !rd(&in.pj[p], 4)) return false;
    std::fclose(f);
    return true;
}

// ---- the plain restatement
static float round_dir(double x, int mode) {
    const int old = std::fegetround();
    std::fesetround(mode);
    volatile double xv = x;
    volatile float f = (float)xv;
    std::fesetround(old);
    return f;
}
struct FBox { float v[6]; };
static FBox fbox(const double* lo, const double* hi) {
    FBox b;
    for (int k = 0; k < 3; ++k) {
        b.v[k] = round_dir(lo[k], FE_DOWNWARD);
        b.v[3 + k] = round_dir(hi[k], FE_UPWARD);
    }
    return b;
}
static bool meets(const FBox& a, const FBox& b) {
    for (int k = 0; k < 3; ++k)
        if (a.v[3 + k] < b.v[k] || a.v[k] > b.v[3 + k]) return false;
    return true;
}
static FBox tri_fbox(const Input& in, int body, int t) {
    const int m = in.mesh[body];
    const int32_t* ix = &in.tri_idx[3 * ((size_t)in.tri_off[m] + t)];
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int c = 0; c < 3; ++c)
        for (int k = 0; k < 3; ++k) {
            const double x = in.wv[3 * (in.wv_off[body] + ix[c]) + k];
            lo[k] = std::min(lo[k], x);
            hi[k] = std::max(hi[k], x);
        }
    return fbox(lo, hi);
}
struct Expect {
    std::vector<int> keptA, keptB;
    std::vector<std::pair<int, int>> hits;
};
static Expect expect(const Input& in, int i, int j) {
    Expect e;
    const FBox bA = fbox(&in.aabb[6 * i], &in.aabb[6 * i + 3]), bB = fbox(&in.aabb[6 * j], &in.aabb[6 * j + 3]);
    std::vector<FBox> fa, fb;
    for (int t = 0; t < in.ntris[in.mesh[i]]; ++t) {
        const FBox f = tri_fbox(in, i, t);
        if (meets(f, bB)) { e.keptA.push_back(t); fa.push_back(f); }
    }
    for (int t = 0; t < in.ntris[in.mesh[j]]; ++t) {
        const FBox f = tri_fbox(in, j, t);
        if (meets(f, bA)) { e.keptB.push_back(t); fb.push_back(f); }
    }
    for (size_t a = 0; a < fa.size(); ++a)
        for (size_t b = 0; b < fb.size(); ++b)
            if (meets(fa[a], fb[b])) e.hits.emplace_back(e.keptA[a], e.keptB[b]);
    return e;
}

// ---- one emulated launch and its check
struct Run {
    int bs, k;
    bool batched, prov;
    int groups;
    bool wave;   // the wave kernel first (else every pair to the workgroup kernel)
    int shards;  // hit-arena shards
};

// q: the wave kernel's queues, the heavy list's counts, the workgroup kernel's queue
template <int BS, int K>
static std::string launch_narrow_emu(const DBodies& B, const DMeshes& M, const DPairs& P, int32_t* li, float* lf,
                                     int32_t tmax, int groups, int use_queue, unsigned* q, bool wave) {
    if (wave) {
        const std::string e = simt::launch((unsigned)groups, NWV * WAVE, (size_t)narrow_wave_lds_bytes(),
                                           [&] { k_narrow_wave<4>(B, M, P, P.heavy, q, use_queue, (int32_t)W_HEAVY_TESTS, nullptr); });
        if (!e.empty()) return "wave kernel: " + e;
    }
    return simt::launch((unsigned)groups, BS, (size_t)narrow_lds_bytes(BS),
                        [&] { k_narrow<BS, K>(B, M, P, li, lf, tmax, use_queue, q + NQ_WG, P.heavy, q + NQ_HEAVY,
                                              q + NQ_HEAVY_BACK, nullptr); });
}

static int check_run(const Input& in, const std::vector<Expect>& ex, const Run& r, long& hits_total) {
    const int np = (int)in.pi.size();
    const int nbody = (int)in.mesh.size();
    int32_t tmax = 1;
    for (int m = 0; m < (int)in.ntris.size(); ++m) tmax = std::max(tmax, in.ntris[m]);
    // device-side arrays, each its own exactly sized heap block
    std::vector<double> wv = in.wv, aabb = in.aabb;
    std::vector<int32_t> nondeg = in.nondeg;
    DBodies B{};
    B.n = nbody; B.mesh = in.mesh.data(); B.wv_off = in.wv_off.data(); B.wv = wv.data(); B.nondeg = nondeg.data();
    B.aabb = aabb.data();
    DMeshes M{};
    M.tri_idx = in.tri_idx.data(); M.nverts = in.nverts.data(); M.tri_off = in.tri_off.data();
    M.ntris = in.ntris.data();
    std::unique_ptr<int32_t[]> order(new int32_t[np]);
    std::unique_ptr<NarrowItem[]> items(new NarrowItem[np]);
    std::vector<int32_t> perm(np);
    std::iota(perm.begin(), perm.end(), 0);
    if (r.batched) std::shuffle(perm.begin(), perm.end(), std::mt19937(12345));
    // the arena: every shard exactly as large as the most loaded one (work
    // item s goes to shard s % shards)
    std::vector<int64_t> load(r.shards, 0);
    for (int s = 0; s < np; ++s) load[s % r.shards] += (int64_t)ex[perm[s]].hits.size();
    const int64_t hit_cap = (int64_t)r.shards * std::max<int64_t>(1, *std::max_element(load.begin(), load.end()));
    for (int s = 0; s < np; ++s) {
        const int p = perm[s], i = in.pi[p], j = in.pj[p];
        order[s] = p;
        NarrowItem& it = items[s];
        std::memset(&it, 0, sizeof it);
        it.p = p; it.i = i; it.j = j;
        it.TA = in.ntris[in.mesh[i]]; it.TB = in.ntris[in.mesh[j]];
        it.VA = in.nverts[in.mesh[i]]; it.VB = in.nverts[in.mesh[j]];
        it.toA = in.tri_off[in.mesh[i]]; it.toB = in.tri_off[in.mesh[j]];
        it.woA = (int32_t)in.wv_off[i]; it.woB = (int32_t)in.wv_off[j];
        it.nondegA = in.nondeg[i];
        box_f(&in.aabb[6 * i], &in.aabb[6 * i + 3], it.boxA);
        box_f(&in.aabb[6 * j], &in.aabb[6 * j + 3], it.boxB);
    }
    std::unique_ptr<int64_t[]> seg_off(new int64_t[np]), tri_checks(new int64_t[np]), tri_tests(new int64_t[np]),
        hit_off(new int64_t[np]);
    std::unique_ptr<int32_t[]> nseg(new int32_t[np]), tri_hits(new int32_t[np]), tri_scanned(new int32_t[np]);
    for (int p = 0; p < np; ++p) { seg_off[p] = nseg[p] = tri_checks[p] = tri_tests[p] = -7; hit_off[p] = -7; tri_hits[p] = tri_scanned[p] = -7; }
    std::unique_ptr<int32_t[]> hits(new int32_t[6 * (size_t)hit_cap + 6]), hit_tri(new int32_t[2 * (size_t)hit_cap + 2]);
    std::fill(hits.get(), hits.get() + 6 * (size_t)hit_cap + 6, -1);
    std::fill(hit_tri.get(), hit_tri.get() + 2 * (size_t)hit_cap + 2, -1);
    int32_t n = np, status = 0;
    unsigned long long hit_top[HIT_SHARDS * HIT_TOP_STRIDE] = {};
    unsigned queue[NQ_COUNTERS] = {};
    std::unique_ptr<int32_t[]> heavy(new int32_t[np]);
    if (!r.wave) {   // every item to the workgroup kernel
        for (int s = 0; s < np; ++s) heavy[s] = s;
        queue[NQ_HEAVY] = (unsigned)np;
    }
    DPairs P{};
    P.cap = np; P.n = &n; P.pi = in.pi.data(); P.pj = in.pj.data();
    P.order = r.batched ? order.get() : nullptr;
    P.items = r.batched ? items.get() : nullptr;
    P.seg_off = seg_off.get(); P.nseg = nseg.get(); P.tri_checks = tri_checks.get(); P.tri_tests = tri_tests.get();
    P.tri_hits = tri_hits.get(); P.tri_scanned = tri_scanned.get(); P.status = &status;
    P.hits = hits.get(); P.hit_tri = r.prov ? hit_tri.get() : nullptr; P.hit_cap = hit_cap; P.hit_top = hit_top;
    P.hit_shards = r.shards;
    P.hit_off = hit_off.get();
    P.heavy = heavy.get();
    std::unique_ptr<int32_t[]> li(new int32_t[(size_t)r.groups * 8 * tmax]);
    std::unique_ptr<float[]> lf(new float[(size_t)r.groups * 12 * tmax]);
    std::string err;
    const int key = r.bs * 100 + r.k;
    const int uq = r.batched ? 1 : 0;
    switch (key) {
        case 6404: err = launch_narrow_emu<64, 4>(B, M, P, li.get(), lf.get(), tmax, r.groups, uq, queue, r.wave); break;
        case 6408: err = launch_narrow_emu<64, 8>(B, M, P, li.get(), lf.get(), tmax, r.groups, uq, queue, r.wave); break;
        case 12804: err = launch_narrow_emu<128, 4>(B, M, P, li.get(), lf.get(), tmax, r.groups, uq, queue, r.wave); break;
        case 12808: err = launch_narrow_emu<128, 8>(B, M, P, li.get(), lf.get(), tmax, r.groups, uq, queue, r.wave); break;
        case 25601: err = launch_narrow_emu<256, 1>(B, M, P, li.get(), lf.get(), tmax, r.groups, uq, queue, r.wave); break;
        case 25602: err = launch_narrow_emu<256, 2>(B, M, P, li.get(), lf.get(), tmax, r.groups, uq, queue, r.wave); break;
        case 25604: err = launch_narrow_emu<256, 4>(B, M, P, li.get(), lf.get(), tmax, r.groups, uq, queue, r.wave); break;
        case 25608: err = launch_narrow_emu<256, 8>(B, M, P, li.get(), lf.get(), tmax, r.groups, uq, queue, r.wave); break;
        case 51202: err = launch_narrow_emu<512, 2>(B, M, P, li.get(), lf.get(), tmax, r.groups, uq, queue, r.wave); break;
        case 51204: err = launch_narrow_emu<512, 4>(B, M, P, li.get(), lf.get(), tmax, r.groups, uq, queue, r.wave); break;
        default: err = "shape not compiled";
    }
    char tag[96];
    std::snprintf(tag, sizeof tag, "%s shape %dx%d %s%s groups %d", r.wave ? "wave+heavy" : "workgroup", r.bs, r.k,
                  r.batched ? "queue+items" : "strided", r.prov ? " prov" : "", r.groups);
    if (!err.empty()) {
        std::printf("FAIL %s: emulation: %s\n", tag, err.c_str());
        return 1;
    }
    int bad = 0;
    auto fail = [&](int p, const char* what, long a, long b) {
        if (bad++ < 12) std::printf("FAIL %s: pair %d: %s: got %ld want %ld\n", tag, p, what, a, b);
    };
    if (status != 0) fail(-1, "status", status, 0);
    int64_t sum = 0;
    std::vector<std::pair<int64_t, int64_t>> ranges;
    for (int p = 0; p < np; ++p) {
        const Expect& e = ex[p];
        const int i = in.pi[p], j = in.pj[p];
        const int TA = in.ntris[in.mesh[i]], TB = in.ntris[in.mesh[j]];
        if (tri_hits[p] != (int32_t)e.hits.size()) { fail(p, "tri_hits", tri_hits[p], (long)e.hits.size()); continue; }
        if (tri_tests[p] != (int64_t)e.keptA.size() * (int64_t)e.keptB.size())
            fail(p, "tri_tests", tri_tests[p], (long)(e.keptA.size() * e.keptB.size()));
        if (tri_scanned[p] != TA + TB) fail(p, "tri_scanned", tri_scanned[p], TA + TB);
        if (tri_checks[p] != (int64_t)in.nondeg[i] * TB) fail(p, "tri_checks", tri_checks[p], (long)in.nondeg[i] * TB);
        if (seg_off[p] != 0 || nseg[p] != 0) fail(p, "seg_off/nseg cleared", seg_off[p], 0);
        sum += tri_hits[p];
        if (tri_hits[p] > 0) ranges.emplace_back(hit_off[p], hit_off[p] + tri_hits[p]);
        const int32_t* ixA = &in.tri_idx[3 * (size_t)in.tri_off[in.mesh[i]]];
        const int32_t* ixB = &in.tri_idx[3 * (size_t)in.tri_off[in.mesh[j]]];
        for (size_t h = 0; h < e.hits.size(); ++h) {
            const int64_t at = hit_off[p] + (int64_t)h;
            if (at < 0 || at >= hit_cap) { fail(p, "hit slot", (long)at, 0); break; }
            const int ta = e.hits[h].first, tb = e.hits[h].second;
            const int32_t want[6] = {(int32_t)(in.wv_off[i] + ixA[3 * ta]), (int32_t)(in.wv_off[i] + ixA[3 * ta + 1]),
                                     (int32_t)(in.wv_off[i] + ixA[3 * ta + 2]), (int32_t)(in.wv_off[j] + ixB[3 * tb]),
                                     (int32_t)(in.wv_off[j] + ixB[3 * tb + 1]), (int32_t)(in.wv_off[j] + ixB[3 * tb + 2])};
            bool ok = true;
            for (int c = 0; c < 6; ++c) ok &= hits[6 * at + c] == want[c];
            if (!ok) { fail(p, "hit record (index in pair)", (long)h, 0); break; }
            if (r.prov && (hit_tri[2 * at] != ta || hit_tri[2 * at + 1] != tb)) {
                fail(p, "hit provenance ta", hit_tri[2 * at], ta);
                break;
            }
        }
    }
    // every pair's range inside the used prefix of its shard, the ranges
    // disjoint, and the shards' counters add up to the hits
    unsigned long long top_sum = 0;
    for (int k = 0; k < r.shards; ++k) top_sum += hit_top[k * HIT_TOP_STRIDE];
    if ((int64_t)top_sum != sum) fail(-1, "shard counters", (long)top_sum, (long)sum);
    const int64_t ssz = hit_cap / r.shards;
    for (const auto& rg : ranges) {
        const int64_t sh = rg.first / ssz;
        if (sh >= r.shards || rg.second > sh * ssz + (int64_t)hit_top[sh * HIT_TOP_STRIDE]) fail(-1, "hit range past its shard", (long)rg.first, 0);
    }
    std::sort(ranges.begin(), ranges.end());
    for (size_t k = 1; k < ranges.size(); ++k)
        if (ranges[k].first < ranges[k - 1].second) fail(-1, "overlapping hit ranges", (long)ranges[k].first, 0);
    hits_total += sum;
    if (bad == 0)
        std::printf("ok %s: %d pairs (%u to the workgroup kernel), %ld hits, %ld barriers, %ld wave operations\n", tag,
                    np, queue[NQ_HEAVY] + queue[NQ_HEAVY_BACK], (long)sum, simt::blk().barriers, simt::blk().waveops);
    simt::blk().barriers = simt::blk().waveops = 0;
    return bad ? 1 : 0;
}

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: narrow_emu INPUT\n");
        return 2;
    }
    Input in;
    if (!read_input(argv[1], in)) {
        std::fprintf(stderr, "narrow_emu: cannot read %s\n", argv[1]);
        return 2;
    }
    const int np = (int)in.pi.size();
    std::vector<Expect> ex(np);
    for (int p = 0; p < np; ++p) ex[p] = expect(in, in.pi[p], in.pj[p]);
    // which of the kernel's paths each shape takes for these pairs
    for (int bs : {64, 128, 256, 512}) {
        int gen = 0, multi = 0, both = 0, bonly = 0, none = 0, vf = 0, direct = 0;
        for (int p = 0; p < np; ++p) {
            const int i = in.pi[p], j = in.pj[p];
            const int VA = in.nverts[in.mesh[i]], VB = in.nverts[in.mesh[j]];
            const int TA = in.ntris[in.mesh[i]], TB = in.ntris[in.mesh[j]];
            const int nA = (int)ex[p].keptA.size(), nB = (int)ex[p].keptB.size();
            if (VA > OC_CAP || VB > OC_CAP) { ++gen; continue; }
            vf += VA + VB <= narrow_vf_cap(bs);
            direct += VA + VB <= narrow_vf_cap(bs) && TA <= 2 * bs && TB <= 2 * bs;
            const bool single = TA <= bs * CULL_PER_THREAD && TB <= bs * CULL_PER_THREAD;
            if (!single) { ++multi; continue; }
            const int kc = narrow_kept_cap(bs);
            if (nB <= kc && nA + nB <= kc) ++both;
            else if (nB <= kc) ++bonly;
            else ++none;
        }
        std::printf("paths at %d threads: general %d, several cull rounds %d, kept in LDS: both %d, B only %d, "
                    "neither %d; vertex boxes in LDS %d, boxes made where tested %d\n", bs, gen, multi, both, bonly,
                    none, vf, direct);
    }
    // which pairs one wave takes (the kernel's own limits)
    int nwave = 0;
    for (int p = 0; p < np; ++p) {
        const int i = in.pi[p], j = in.pj[p];
        const int VA = in.nverts[in.mesh[i]], VB = in.nverts[in.mesh[j]];
        const int nA = (int)ex[p].keptA.size(), nB = (int)ex[p].keptB.size();
        const int ocb = (VA + VB + 15) & ~15;
        nwave += VA + VB <= W_OC_MAX && nA + nB <= (WSLICE - ocb) / 40 && (int64_t)nA * nB <= W_HEAVY_TESTS;
    }
    std::printf("wave kernel takes %d of %d pairs\n", nwave, np);
    const Run runs[] = {
        {64, 4, true, false, 3, false, 1},  {64, 8, false, true, 5, false, 7},  {128, 4, true, true, 2, false, 64},
        {128, 8, false, false, 4, false, 3}, {256, 1, true, false, 2, false, 1}, {256, 2, false, false, 3, false, 64},
        {256, 4, true, true, 3, false, 5},  {256, 8, false, true, 2, false, 64}, {512, 2, true, false, 2, false, 2},
        {512, 4, true, true, 2, false, 64}, {512, 4, false, false, 3, false, 1},
        {512, 4, true, false, 3, true, 64}, {512, 4, false, true, 2, true, 13}, {256, 4, true, true, 2, true, 1},
        {64, 4, false, false, 4, true, 64},
    };
    int fails = 0;
    long hits_total = 0;
    for (const Run& r : runs) fails += check_run(in, ex, r, hits_total);
    if (fails) {
        std::printf("narrow_emu: %d runs failed\n", fails);
        return 1;
    }
    std::printf("narrow_emu ok: %d runs\n", (int)(sizeof(runs) / sizeof(runs[0])));
    return 0;
}
