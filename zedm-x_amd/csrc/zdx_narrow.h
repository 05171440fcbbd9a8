// zdx_narrow.h — narrow phase, part 1: the cull-and-sweep kernel k_narrow
// and its float-box helpers.
//
// Textually included inside namespace zdx::{anonymous} after zdx_wave.h: by
// zdx_kernels.hip for gfx950, and by the host SIMT emulation of the kernel
// (tests/csrc/narrow_emu.cpp), which runs this same source thread by thread
// (each workgroup's threads as fibers, real barriers, wave operations that
// require every lane of the wave at the same call) and checks its hit lists
// against a plain restatement of the test.  ZDX_DYN_LDS declares the
// workgroup's dynamic LDS.  No includes or namespaces of its own.

#ifdef ZDX_WALK_PROFILE
// per work item: nA, nB, hits, p, TA, TB, VA, VB, then ticks of 8 phases (item
// read, vertex pass, cull tests, kept boxes, sweep count, sweep emit, tail),
// then the workgroup (or wave) that took it, its start and end clocks and the
// kernel (0 workgroup kernel, 1 wave kernel)
constexpr int NP_REC = 20;
__device__ long long g_np_pairs[NP_REC * 16384];
#define NP_START() unsigned long long np_t0 = wall_clock64(); unsigned np_ph[8] = {0, 0, 0, 0, 0, 0, 0, 0}
#define NP_MARK(k)                                                      \
    do {                                                                \
        const unsigned long long t_ = wall_clock64();                   \
        np_ph[(k) & 7] += (unsigned)(t_ - np_t0);                       \
        np_t0 = t_;                                                     \
    } while (0)
#else
#define NP_MARK(k) do {} while (0)
#define NP_START() do {} while (0)
#endif

// outward-rounded float AABB of a double box
__device__ inline void box_f(const double* lo, const double* hi, float* f) {
    f[0] = __double2float_rd(lo[0]); f[1] = __double2float_rd(lo[1]); f[2] = __double2float_rd(lo[2]);
    f[3] = __double2float_ru(hi[0]); f[4] = __double2float_ru(hi[1]); f[5] = __double2float_ru(hi[2]);
}
__device__ inline bool ovl_f(const float* a, const float* b) {
    return !(a[3] < b[0] || a[0] > b[3] || a[4] < b[1] || a[1] > b[4] || a[5] < b[2] || a[2] > b[5]);
}
__device__ inline void tri_box(const V3* t, double* lo, double* hi) {
    lo[0] = min3(t[0].x, t[1].x, t[2].x); lo[1] = min3(t[0].y, t[1].y, t[2].y); lo[2] = min3(t[0].z, t[1].z, t[2].z);
    hi[0] = max3(t[0].x, t[1].x, t[2].x); hi[1] = max3(t[0].y, t[1].y, t[2].y); hi[2] = max3(t[0].z, t[1].z, t[2].z);
}
__device__ inline void load_tri(const double* wv, const int32_t* ti, int t, V3* out) {
    out[0] = ld3(wv + 3 * ti[3 * t]);
    out[1] = ld3(wv + 3 * ti[3 * t + 1]);
    out[2] = ld3(wv + 3 * ti[3 * t + 2]);
}

// Vertex outcodes against a float box: bit 2k set when the vertex lies below
// the box on axis k with its coordinate rounded up, bit 2k+1 when it lies
// above with its coordinate rounded down.  A triangle's outward-rounded float
// AABB misses the box exactly when the three outcodes share a set bit
// (rounding is monotone, so rounding the vertex coordinates commutes with the
// min/max of the box), which makes the cull below the same test as
// ovl_f(box_f(tri_box(...)), box) with one load per vertex instead of three
// per triangle.  The codes come from six double compares: rounding x up gives
// a float below the float L exactly when x <= the largest float below L, and
// rounding down gives one above H exactly when x >= the smallest float above
// H, so a box gives six double thresholds once and a vertex needs no
// conversion.
struct OcThr { double lo[3], hi[3]; };
__device__ inline OcThr oc_thresholds(const float* bx) {
    OcThr o;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        o.lo[k] = (double)nextafterf(bx[k], -INFINITY);
        o.hi[k] = (double)nextafterf(bx[3 + k], INFINITY);
    }
    return o;
}
__device__ inline unsigned vertex_outcode(const double* w, const OcThr& o) {
    return (unsigned)(w[0] <= o.lo[0]) | ((unsigned)(w[0] >= o.hi[0]) << 1) | ((unsigned)(w[1] <= o.lo[1]) << 2) |
           ((unsigned)(w[1] >= o.hi[1]) << 3) | ((unsigned)(w[2] <= o.lo[2]) << 4) | ((unsigned)(w[2] >= o.hi[2]) << 5);
}

// a triangle's three vertex indices, one 12-byte load
struct Tri3 { int32_t a, b, c; };
__device__ inline Tri3 ld_tri(const int32_t* ti, int t) { return ((const Tri3*)ti)[t]; }

constexpr int OC_CAP = 2048;   // vertices per body whose outcodes fit in LDS
constexpr int CULL_PER_THREAD = 8;

// A work item's share of the hit arena: `hits` slots in the item's shard,
// returned as the first slot, or ~0 (and ST_HITS raised) when the shard is
// full -- its counter still records what was asked for, which sizes the
// arena for the step's retry.
__device__ inline unsigned long long reserve_hits(const DPairs& P, int item, int hits) {
    if (hits <= 0) return 0ull;
    const int sh = item % P.hit_shards;
    const unsigned long long size = (unsigned long long)(P.hit_cap / P.hit_shards);
    const unsigned long long hb = atomicAdd(P.hit_top + sh * HIT_TOP_STRIDE, (unsigned long long)hits);
    if (hb + (unsigned long long)hits > size) {
        atomicOr(P.status, (int)ST_HITS);
        return ~0ull;
    }
    return (unsigned long long)sh * size + hb;
}

// The float-box tests of the range [g, g + n) of the flattened (ca, cb)
// product of nA x nB kept triangles, row by row: f(k, ca, cb) for each pair
// of boxes that overlap, k = index within the range.  The K box loads of a
// row segment are independent of each other (the index is clamped into the
// row, so they are always in bounds), which lets them be in flight together.
template <int K, class F>
__device__ inline void sweep_range(int64_t g, int n, int nB, const float* FAx, const float* FBx, F&& f) {
    if (n <= 0) return;
    int ca, cb;
    if (g <= 0x7fffffff) { ca = (int)g / nB; cb = (int)g - ca * nB; }
    else { ca = (int)(g / nB); cb = (int)(g - (int64_t)ca * nB); }
    int k = 0;
    while (k < n) {
        float fa[6];
#pragma unroll
        for (int q = 0; q < 6; ++q) fa[q] = FAx[6 * ca + q];
        const int len = nB - cb < n - k ? nB - cb : n - k;
        for (int u = 0; u < len; u += K) {
            float fb[K][6];
#pragma unroll
            for (int j = 0; j < K; ++j) {
                const int c = cb + (u + j < len ? u + j : len - 1);
#pragma unroll
                for (int q = 0; q < 6; ++q) fb[j][q] = FBx[6 * c + q];
            }
#pragma unroll
            for (int j = 0; j < K; ++j)
                if (u + j < len && ovl_f(fa, fb[j])) f(k + u + j, ca, cb + u + j);
        }
        k += len;
        ++ca;
        cb = 0;
    }
}

// LDS of the narrow kernel per workgroup size: the rounded coordinates of up
// to narrow_vf_cap vertices (both bodies, 24 bytes each) and up to
// narrow_kept_cap kept triangles (box 24 bytes, world-vertex indices 12,
// triangle id 4), sized so that 2 workgroups of 512, 5 of 256, 7 of 128 or 12
// of 64 threads share a CU's 160 KB with the static arrays (about 5.3 KB).
__host__ __device__ constexpr int narrow_vf_cap(int bs) { return bs >= 512 ? 1024 : (bs >= 256 ? 512 : 2 * bs); }
__host__ __device__ constexpr int narrow_kept_cap(int bs) { return bs >= 512 ? 1280 : (bs >= 256 ? 352 : 2 * bs); }
__host__ __device__ constexpr int narrow_lds_bytes(int bs) { return 24 * narrow_vf_cap(bs) + 40 * narrow_kept_cap(bs); }
static_assert(2 * (narrow_lds_bytes(512) + 5400) <= 160 * 1024, "two 512-thread narrow workgroups per CU");
static_assert(5 * (narrow_lds_bytes(256) + 5400) <= 160 * 1024, "five 256-thread narrow workgroups per CU");

// Narrow phase, part 1 (cull and sweep): persistent workgroups, one pair at a
// time.  Each side's triangles are culled against the other body's AABB, then
// the flattened (ta, tb) product of the survivors is swept with the float
// AABB test; the triangle pairs that pass ("hits") are written in (ta, tb)
// order to an exactly sized slice of the hit arena, so that every pair's
// hits are contiguous.  Pairs with hits are listed for part 2
// (k_narrow_exact).  Register budget: 5 waves per SIMD for 256 threads (5
// workgroups per CU), 4 for 512 threads (the same 2 workgroups per CU as 5
// would give, with room to keep everything in registers).
//
// Bodies of at most OC_CAP vertices each (the common case) take one load per
// vertex: its outcode against the other body's box and, when both bodies'
// vertices fit, its coordinates rounded down and up go to LDS; each kept
// triangle's float box is then the min / max of its vertices' rounded
// coordinates (rounding is monotone, so this is box_f(tri_box(...))), written
// with its world-vertex indices at its rank in triangle order by the thread
// that tested it.  The kept data of both sides stays in LDS for the sweep
// when it fits (B's alone when only B's fits), else it goes to the
// workgroup's global lists.  The next work item is read while the current
// one is swept.
template <int BS, int K>
__global__ void __launch_bounds__(BS, BS >= 512 ? 4 : (BS >= 256 ? 5 : 1)) k_narrow(DBodies B, DMeshes M, DPairs P, int32_t* list_i, float* list_f,
                                               int32_t tmax, int32_t use_queue, unsigned int* queue,
                                               const int32_t* hlist, const unsigned* hcount, const unsigned* hback,
                                               unsigned long long* stamp) {
    stamp_here(stamp);
    constexpr int VF = narrow_vf_cap(BS), KC = narrow_kept_cap(BS);
    ZDX_DYN_LDS(s_nb);
    float* s_vf = (float*)s_nb;                    // 6 x VF: coordinates rounded down (x, y, z), then up
    float* s_kf = s_vf + 6 * VF;                   // kept boxes: B's, then A's
    int32_t* s_kv = (int32_t*)(s_kf + 6 * KC);     // kept world-vertex indices, same order
    int32_t* s_kt = s_kv + 3 * KC;                 // kept triangle ids, same order
    __shared__ int s_wave[BS / WAVE];
    __shared__ int s_next;
    __shared__ NarrowItem s_it;
    __shared__ unsigned char s_oc[2 * OC_CAP];
    __shared__ unsigned long long s_kmask[2 * WAVE];   // cull: keep ballots per (layer, wave), A then B
    __shared__ unsigned long long s_base;
    static_assert(sizeof(NarrowItem) == 32 * sizeof(uint32_t), "a work item is 32 words");
    const int tid = threadIdx.x;
    int32_t* LA = list_i + (int64_t)blockIdx.x * 8 * tmax;
    int32_t* LB = LA + tmax;
    int32_t* VA3 = LA + 2 * (int64_t)tmax;
    int32_t* VB3 = LA + 5 * (int64_t)tmax;
    float* FA = list_f + (int64_t)blockIdx.x * 12 * tmax;
    float* FBg = FA + 6 * (int64_t)tmax;
    // the work: every pair, or (with hlist) the pairs listed there: *hcount
    // from the front of hlist, then *hback from its end backwards (the item
    // count is the list's length)
    const int hfront = hlist ? (int)*hcount : 0, hend = *P.n - 1;
    const int nq = hlist ? hfront + (int)*hback : *P.n;
    auto listed = [&](int k) { return k < hfront ? hlist[k] : hlist[hend - (k - hfront)]; };
    // first queue entry; with the work queue each workgroup asks for its next
    // entry while it works on the current one
    if (tid == 0) s_next = use_queue ? (int)atomicAdd(queue, 1u) : (int)blockIdx.x;
    __syncthreads();
    int qi = uni(s_next);
    int item = qi < nq ? (hlist ? uni(listed(qi)) : qi) : 0;
    if (P.items && tid < 32 && qi < nq) ((uint32_t*)&s_it)[tid] = ((const uint32_t*)(P.items + item))[tid];
    __syncthreads();
    NP_START();
    while (qi < nq) {
        int next = 0;
        if (tid == 0) next = use_queue ? (int)atomicAdd(queue, 1u) : qi + (int)gridDim.x;
        int p, i, j, TA, TB, VA, VB, ndA;
        int64_t oA, oB;
        const int32_t* tiA;
        const int32_t* tiB;
        float boxA[6], boxB[6];
        if (P.items) {
            // the item's fields are the workgroup's: kept in scalar registers
            const NarrowItem& it = s_it;
            p = uni(it.p); i = uni(it.i); j = uni(it.j);
            TA = uni(it.TA); TB = uni(it.TB); VA = uni(it.VA); VB = uni(it.VB); ndA = uni(it.nondegA);
            tiA = M.tri_idx + 3 * (int64_t)uni(it.toA);
            tiB = M.tri_idx + 3 * (int64_t)uni(it.toB);
            oA = uni(it.woA); oB = uni(it.woB);
            for (int k = 0; k < 6; ++k) { boxA[k] = uni(it.boxA[k]); boxB[k] = uni(it.boxB[k]); }
        } else {
            p = P.order ? P.order[item] : item;
            i = P.pi[p]; j = P.pj[p];
            const int mA = B.mesh[i], mB = B.mesh[j];
            TA = M.ntris[mA]; TB = M.ntris[mB];
            VA = M.nverts[mA]; VB = M.nverts[mB];
            ndA = B.nondeg[i];
            tiA = M.tri_idx + 3 * (int64_t)M.tri_off[mA];
            tiB = M.tri_idx + 3 * (int64_t)M.tri_off[mB];
            oA = B.wv_off[i]; oB = B.wv_off[j];
            box_f(B.aabb + 6 * (int64_t)i, B.aabb + 6 * (int64_t)i + 3, boxA);
            box_f(B.aabb + 6 * (int64_t)j, B.aabb + 6 * (int64_t)j + 3, boxB);
        }
        const double* wvA = B.wv + 3 * oA;
        const double* wvB = B.wv + 3 * oB;
        NP_MARK(0);
        // ---- cull each side against the other body's AABB (conservative)
        int nA = 0, nB = 0;
        bool kla = false, klb = false;   // A's / B's kept data in LDS
        if (VA <= OC_CAP && VB <= OC_CAP) {
            constexpr int CPT = CULL_PER_THREAD;
            constexpr int NW = BS / WAVE;
            static_assert(CPT * NW <= WAVE, "keep masks of one round fit one wave");
            const int wid = tid / WAVE, lane = tid % WAVE;
            const uint64_t below = (1ull << lane) - 1ull;
            const bool single = TA <= BS * CPT && TB <= BS * CPT;
            const bool vfl = VA + VB <= VF;
            // direct: each kept triangle's box is made by the thread that
            // tested it, from registers and LDS alone (both bodies' rounded
            // coordinates fit the LDS, and each side has at most two layers of
            // triangles, whose indices are loaded before the vertex pass);
            // otherwise the kept ids go to their ranks and one parallel pass
            // makes the boxes
            const bool direct = vfl && TA <= 2 * BS && TB <= 2 * BS;
            int ra[2][3], rb[2][3];
            if (direct) {
#pragma unroll
                for (int c = 0; c < 2; ++c) {
                    const int t = c * BS + tid;
                    if (t < TA) { const Tri3 q = ld_tri(tiA, t); ra[c][0] = q.a; ra[c][1] = q.b; ra[c][2] = q.c; }
                    if (t < TB) { const Tri3 q = ld_tri(tiB, t); rb[c][0] = q.a; rb[c][1] = q.b; rb[c][2] = q.c; }
                }
            }
            // outcodes (and rounded coordinates) of each body's vertices against
            // the other body's box
            constexpr int UV = BS <= 64 ? 8 : 4;   // vertices per thread per round, loads in flight
            auto vpass = [&](const double* wv, int n, int voff, const float* other, unsigned char* ocs) {
                const OcThr th = oc_thresholds(other);
                for (int v0 = 0; v0 < n; v0 += UV * BS) {
                    double w[UV][3];
#pragma unroll
                    for (int u = 0; u < UV; ++u) {   // all loads first, then the codes
                        const int v = v0 + u * BS + tid;
                        if (v < n) { w[u][0] = wv[3 * v]; w[u][1] = wv[3 * v + 1]; w[u][2] = wv[3 * v + 2]; }
                    }
#pragma unroll
                    for (int u = 0; u < UV; ++u) {
                        const int v = v0 + u * BS + tid;
                        if (v < n) {
                            ocs[v] = (unsigned char)vertex_outcode(w[u], th);
                            if (vfl) {
#pragma unroll
                                for (int k = 0; k < 3; ++k) {
                                    s_vf[k * VF + voff + v] = __double2float_rd(w[u][k]);
                                    s_vf[(3 + k) * VF + voff + v] = __double2float_ru(w[u][k]);
                                }
                            }
                        }
                    }
                }
            };
            vpass(wvA, VA, 0, boxB, s_oc);
            vpass(wvB, VB, VA, boxA, s_oc + OC_CAP);
            __syncthreads();
            NP_MARK(1);
            // a triangle's float box: the min / max of its vertices' rounded
            // coordinates (LDS), or box_f of its vertices' box (global)
            auto tri_fbox = [&](bool a, int i0, int i1, int i2, float* f) {
                if (vfl) {
                    const int s = a ? 0 : VA;
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        const float* lo = s_vf + k * VF + s;
                        const float* hi = s_vf + (3 + k) * VF + s;
                        float m = lo[i0], x = hi[i0];
                        const float l1 = lo[i1], l2 = lo[i2], h1 = hi[i1], h2 = hi[i2];
                        if (l1 < m) m = l1;
                        if (l2 < m) m = l2;
                        if (x < h1) x = h1;
                        if (x < h2) x = h2;
                        f[k] = m;
                        f[3 + k] = x;
                    }
                } else {
                    const double* wv = a ? wvA : wvB;
                    V3 tr[3] = {ld3(wv + 3 * i0), ld3(wv + 3 * i1), ld3(wv + 3 * i2)};
                    double lo[3], hi[3];
                    tri_box(tr, lo, hi);
                    box_f(lo, hi, f);
                }
            };
            auto put = [&](bool a, int i0, int i1, int i2, const float* f, float* F, int32_t* V3i) {
#pragma unroll
                for (int k = 0; k < 6; ++k) F[k] = f[k];
                const int64_t o = a ? oA : oB;
                V3i[0] = (int32_t)(o + i0); V3i[1] = (int32_t)(o + i1); V3i[2] = (int32_t)(o + i2);
            };
            // CPT layers of BS consecutive triangles per round (lane-consecutive
            // index loads); the keep ballots of every (layer, wave) go to LDS
            // and each wave turns their popcounts into ranks in triangle order
            for (int base = 0; base < TA || base < TB; base += BS * CPT) {
                unsigned ka = 0u, kb = 0u;
#pragma unroll
                for (int c = 0; c < CPT; ++c) {
                    const int t = base + c * BS + tid;
                    if (t < TA) {
                        int i0, i1, i2;
                        if (direct && c < 2) { i0 = ra[c < 2 ? c : 0][0]; i1 = ra[c < 2 ? c : 0][1]; i2 = ra[c < 2 ? c : 0][2]; }
                        else { const Tri3 q = ld_tri(tiA, t); i0 = q.a; i1 = q.b; i2 = q.c; }
                        if ((s_oc[i0] & s_oc[i1] & s_oc[i2]) == 0) ka |= 1u << c;
                    }
                    if (t < TB) {
                        int i0, i1, i2;
                        if (direct && c < 2) { i0 = rb[c < 2 ? c : 0][0]; i1 = rb[c < 2 ? c : 0][1]; i2 = rb[c < 2 ? c : 0][2]; }
                        else { const Tri3 q = ld_tri(tiB, t); i0 = q.a; i1 = q.b; i2 = q.c; }
                        if ((s_oc[OC_CAP + i0] & s_oc[OC_CAP + i1] & s_oc[OC_CAP + i2]) == 0) kb |= 1u << c;
                    }
                }
#pragma unroll
                for (int c = 0; c < CPT; ++c) {
                    const uint64_t ma = __ballot((ka >> c) & 1u), mb = __ballot((kb >> c) & 1u);
                    if (lane == 0) { s_kmask[c * NW + wid] = ma; s_kmask[WAVE + c * NW + wid] = mb; }
                }
                __syncthreads();
                // inclusive prefix of the popcounts in (layer, wave) order, per side
                const int q = lane;
                int pa = q < CPT * NW ? __popcll(s_kmask[q]) : 0;
                int pb = q < CPT * NW ? __popcll(s_kmask[WAVE + q]) : 0;
#pragma unroll
                for (int d = 1; d < WAVE; d <<= 1) {
                    const int ya = __shfl_up(pa, d, WAVE), yb = __shfl_up(pb, d, WAVE);
                    if (lane >= d) { pa += ya; pb += yb; }
                }
                const int tot_a = uni(__shfl(pa, CPT * NW - 1, WAVE)), tot_b = uni(__shfl(pb, CPT * NW - 1, WAVE));
                NP_MARK(2);
                if (base == 0) {   // where the kept data goes (one round: all counts known here)
                    klb = single && tot_b <= KC;
                    kla = klb && tot_a + tot_b <= KC;
                }
                // (with one round, tot_b is B's final count)
                int32_t* TAx = kla ? s_kt + tot_b : LA;
                int32_t* TBx = klb ? s_kt : LB;
                if (direct) {
                    float* FAx = kla ? s_kf + 6 * tot_b : FA;
                    float* FBx = klb ? s_kf : FBg;
                    int32_t* VAx = kla ? s_kv + 3 * tot_b : VA3;
                    int32_t* VBx = klb ? s_kv : VB3;
                    // one layer at a time (unrolled, every layer's LDS loads
                    // would be held in registers at once)
#pragma unroll 1
                    for (int c = 0; c < 2; ++c) {
                        const int slot = c * NW + wid;
                        const int t = c * BS + tid;
                        // the (layer, wave) prefixes, read with every lane active
                        // (a lane shuffle from an inactive lane returns 0)
                        const int ea = __shfl(pa, slot, WAVE) - __popcll(s_kmask[slot]);
                        const int eb = __shfl(pb, slot, WAVE) - __popcll(s_kmask[WAVE + slot]);
                        float f[6];
                        if ((ka >> c) & 1u) {
                            const int r = ea + __popcll(s_kmask[slot] & below);
                            const int i0 = c == 0 ? ra[0][0] : ra[1][0], i1 = c == 0 ? ra[0][1] : ra[1][1],
                                      i2 = c == 0 ? ra[0][2] : ra[1][2];
                            tri_fbox(true, i0, i1, i2, f);
                            put(true, i0, i1, i2, f, FAx + 6 * r, VAx + 3 * r);
                            TAx[r] = t;
                        }
                        if ((kb >> c) & 1u) {
                            const int r = eb + __popcll(s_kmask[WAVE + slot] & below);
                            const int i0 = c == 0 ? rb[0][0] : rb[1][0], i1 = c == 0 ? rb[0][1] : rb[1][1],
                                      i2 = c == 0 ? rb[0][2] : rb[1][2];
                            tri_fbox(false, i0, i1, i2, f);
                            put(false, i0, i1, i2, f, FBx + 6 * r, VBx + 3 * r);
                            TBx[r] = t;
                        }
                    }
                } else {
#pragma unroll
                    for (int c = 0; c < CPT; ++c) {
                        const int slot = c * NW + wid;
                        const int t = base + c * BS + tid;
                        const int ea = __shfl(pa, slot, WAVE) - __popcll(s_kmask[slot]);
                        const int eb = __shfl(pb, slot, WAVE) - __popcll(s_kmask[WAVE + slot]);
                        if ((ka >> c) & 1u) TAx[nA + ea + __popcll(s_kmask[slot] & below)] = t;
                        if ((kb >> c) & 1u) TBx[nB + eb + __popcll(s_kmask[WAVE + slot] & below)] = t;
                    }
                }
                __syncthreads();   // the masks are rewritten by the next round
                NP_MARK(3);
                nA += tot_a;
                nB += tot_b;
            }
            if (!direct) {
                // the kept triangles' boxes and world-vertex indices, all lanes
                // in parallel
                const int32_t* TAx = kla ? s_kt + nB : LA;
                const int32_t* TBx = klb ? s_kt : LB;
                float* FAx = kla ? s_kf + 6 * nB : FA;
                float* FBx = klb ? s_kf : FBg;
                int32_t* VAx = kla ? s_kv + 3 * nB : VA3;
                int32_t* VBx = klb ? s_kv : VB3;
                for (int k = tid; k < nA + nB; k += BS) {
                    const bool a = k < nA;
                    const int q = a ? k : k - nA;
                    const int t = a ? TAx[q] : TBx[q];
                    const int32_t* ix = (a ? tiA : tiB) + 3 * t;
                    const int i0 = ix[0], i1 = ix[1], i2 = ix[2];
                    float f[6];
                    tri_fbox(a, i0, i1, i2, f);
                    put(a, i0, i1, i2, f, a ? FAx + 6 * q : FBx + 6 * q, a ? VAx + 3 * q : VBx + 3 * q);
                }
                __syncthreads();
                NP_MARK(3);
            }
        } else {
            for (int base = 0; base < TA; base += BS) {
                int t = base + tid;
                bool keep = false;
                float f[6];
                if (t < TA) {
                    V3 tr[3]; double lo[3], hi[3];
                    load_tri(wvA, tiA, t, tr);
                    tri_box(tr, lo, hi);
                    box_f(lo, hi, f);
                    keep = ovl_f(f, boxB);
                }
                int tot;
                int r = block_rank<BS>(keep, s_wave, tot);
                if (keep) LA[nA + r] = t;
                nA += tot;
            }
            NP_MARK(1);
            for (int base = 0; base < TB; base += BS) {
                int t = base + tid;
                bool keep = false;
                float f[6];
                if (t < TB) {
                    V3 tr[3]; double lo[3], hi[3];
                    load_tri(wvB, tiB, t, tr);
                    tri_box(tr, lo, hi);
                    box_f(lo, hi, f);
                    keep = ovl_f(f, boxA);
                }
                int tot;
                int r = block_rank<BS>(keep, s_wave, tot);
                if (keep) LB[nB + r] = t;
                nB += tot;
            }
            __syncthreads();
            NP_MARK(2);
            // float boxes and world-vertex indices of the kept triangles, all
            // lanes in parallel; B's boxes in LDS when they fit
            klb = nB <= KC;
            float* FBx = klb ? s_kf : FBg;
            for (int k = tid; k < nA + nB; k += BS) {
                const bool a = k < nA;
                const int q = a ? k : k - nA;
                const int t = a ? LA[q] : LB[q];
                const int32_t* ix = (a ? tiA : tiB) + 3 * t;
                const int i0 = ix[0], i1 = ix[1], i2 = ix[2];
                const double* wv = a ? wvA : wvB;
                V3 tr[3] = {ld3(wv + 3 * i0), ld3(wv + 3 * i1), ld3(wv + 3 * i2)};
                double lo[3], hi[3];
                tri_box(tr, lo, hi);
                box_f(lo, hi, a ? FA + 6 * q : FBx + 6 * q);
                const int64_t o = a ? oA : oB;
                int32_t* v3i = (a ? VA3 : (klb ? s_kv : VB3)) + 3 * q;
                v3i[0] = (int32_t)(o + i0); v3i[1] = (int32_t)(o + i1); v3i[2] = (int32_t)(o + i2);
                if (!a && klb) s_kt[q] = t;
            }
        }
        if (tid == 0) s_next = next;
        __syncthreads();
        NP_MARK(3);
        // the next work item, read now and stored to LDS after the sweep
        uint32_t nxt_word = 0u;
        const int nxt = uni(s_next);
        const int nxt_item = nxt < nq ? (hlist ? listed(nxt) : nxt) : 0;
        if (P.items && tid < 32 && nxt < nq) nxt_word = ((const uint32_t*)(P.items + nxt_item))[tid];
        // ---- ordered sweep of the culled product into the hit arena.  Each
        // thread takes one contiguous range of the flattened (ta, tb) product,
        // so the product's order is the threads' order: a counting pass, one
        // block scan, an exactly sized range of the arena, then the write pass
        // (from the hit bits of the counting pass when the range has at most 64
        // items, else by testing again).  No barrier inside the sweep.
        const int64_t total = (int64_t)nA * nB;
        const int64_t per = (total + BS - 1) / BS;
        const int64_t g_begin = total < (int64_t)tid * per ? total : (int64_t)tid * per;
        const int64_t g_end = total < g_begin + per ? total : g_begin + per;
        auto split = [&](int64_t g, int& ca, int& cb) {
            if (total <= 0x7fffffff) { ca = (int)g / nB; cb = (int)g - ca * nB; }
            else { ca = (int)(g / nB); cb = (int)(g - (int64_t)ca * nB); }
        };
        // the sweep with B's boxes at FBx (LDS or global, one instance each so
        // that the inner loop's loads are typed by address space); A's kept
        // data at FAx, VAx, TAx, B's others at VBx, TBx
        auto sweep = [&](const float* FBx, const float* FAx, const int32_t* VAx, const int32_t* VBx,
                         const int32_t* TAx, const int32_t* TBx) __attribute__((always_inline)) -> int {
            // f(k, ca, cb) for each hit of the range, k = item index within the range
            auto for_hits = [&](auto&& f) { sweep_range<K>(g_begin, (int)(g_end - g_begin), nB, FAx, FBx, f); };
            // a hit record: the six world-vertex indices of (ta, tb); with
            // provenance also (ta, tb) itself
            auto emit = [&](int64_t at, int ca, int cb) {
                int2* dst = (int2*)(P.hits + 6 * at);
                const int32_t* va = VAx + 3 * ca;
                const int32_t* vb = VBx + 3 * cb;
                dst[0] = make_int2(va[0], va[1]);
                dst[1] = make_int2(va[2], vb[0]);
                dst[2] = make_int2(vb[1], vb[2]);
                if (P.hit_tri) ((int2*)P.hit_tri)[at] = make_int2(TAx[ca], TBx[cb]);
            };
            int cnt = 0;
            uint64_t bits = 0ull;   // hit bits of the range when per <= 64
            for_hits([&](int k, int, int) {
                ++cnt;
                if (k < 64) bits |= 1ull << k;
            });
            int hits;
            const int off = block_excl<BS>(cnt, s_wave, hits);
            if (tid == 0) s_base = reserve_hits(P, item, hits);
            __syncthreads();
            NP_MARK(4);
            const unsigned long long hbase = s_base;
            if (hbase != ~0ull && cnt > 0) {
                int64_t at = (int64_t)hbase + off;
                if (per <= 64) {
                    for (uint64_t m = bits; m; m &= m - 1) {
                        int ca, cb;
                        split(g_begin + ctz64(m), ca, cb);
                        emit(at++, ca, cb);
                    }
                } else {
                    for_hits([&](int, int ca, int cb) { emit(at++, ca, cb); });
                }
            }
            return hits;
        };
        const float* FAx = kla ? s_kf + 6 * nB : FA;
        const int32_t* VAx = kla ? s_kv + 3 * nB : VA3;
        const int32_t* TAx = kla ? s_kt + nB : LA;
        const int hits = klb ? sweep(s_kf, FAx, VAx, s_kv, TAx, s_kt) : sweep(FBg, FA, VA3, VB3, LA, LB);
        NP_MARK(5);
        if (tid == 0) {
            const bool ok = s_base != ~0ull;
            P.hit_off[p] = ok ? (int64_t)s_base : 0;
            P.tri_hits[p] = ok ? hits : 0;
            P.tri_checks[p] = (int64_t)ndA * TB;
            P.tri_tests[p] = total;
            P.tri_scanned[p] = TA + TB;
            P.seg_off[p] = 0;   // set by the compaction for pairs with hits
            P.nseg[p] = 0;
        }
        if (P.items && tid < 32) ((uint32_t*)&s_it)[tid] = nxt_word;
        __syncthreads();
        NP_MARK(6);
#ifdef ZDX_WALK_PROFILE
        if (tid == 0 && item < 16384) {
            long long* r = g_np_pairs + NP_REC * (int64_t)item;
            r[0] = nA; r[1] = nB; r[2] = hits; r[3] = p;
            r[4] = TA; r[5] = TB; r[6] = VA; r[7] = VB;
            long long tot = 0;
            for (int k = 0; k < 8; ++k) { r[8 + k] = np_ph[k]; tot += np_ph[k]; np_ph[k] = 0; }
            r[16] = blockIdx.x; r[17] = (long long)np_t0 - tot; r[18] = (long long)np_t0; r[19] = 0;
        }
#endif
        qi = nxt;
        item = uni(nxt_item);
    }
}

// ---- one wave per pair
// Waves of the wave kernel per workgroup, LDS per wave, and the limits of a
// pair that one wave takes: its two bodies' vertices (one outcode byte each)
// and kept triangles (id 4 bytes, float box 24, world-vertex indices 12) fit
// the wave's LDS, and its float-box tests number at most the launch's
// heavy_tests (W_HEAVY_TESTS unless tuned: fewer leave more pairs to the
// workgroup kernel, which shortens the wave kernel's tail when pairs are few;
// more keep more pairs on the cheaper wave path).
// Other pairs ("heavy") are listed for the workgroup kernel.  4 workgroups of
// 4 waves share a CU's 160 KB.
constexpr int NWV = 4;
constexpr int WSLICE = 9728;
// the narrow phase's counters (zeroed before it), each on a 128-byte line of
// its own (atomics on one line serialise): WQ work queues of the wave kernel,
// then the heavy list's two counts (front, back) and the workgroup kernel's queue
constexpr int WQ = 64;
constexpr int QSTRIDE = 32;   // in 4-byte counters
constexpr int NQ_HEAVY = WQ * QSTRIDE, NQ_WG = (WQ + 1) * QSTRIDE, NQ_HEAVY_BACK = (WQ + 2) * QSTRIDE,
              NQ_COUNTERS = (WQ + 3) * QSTRIDE;
// heavy pairs of at least this many float-box tests (or too many vertices for
// the wave) go to the front of the heavy list, the others to its back, so that
// the workgroup kernel starts the costliest first
constexpr int W_HEAVY_FIRST = 16384;
constexpr int W_OC_MAX = 2048;
constexpr int64_t W_HEAVY_TESTS = 256 * WAVE;
__host__ __device__ constexpr int narrow_wave_lds_bytes() { return NWV * WSLICE; }
static_assert(4 * (NWV * WSLICE + 64) <= 160 * 1024, "four wave-kernel workgroups per CU");

// LDS written by some lanes of a wave, read by others: the wave's LDS
// operations complete in order, so this only keeps the compiler from moving
// accesses across it.
__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Narrow phase, part 1, one wave per pair: the same cull, sweep and hit
// records as k_narrow, with no workgroup barrier.  Each wave takes work items
// from one of (up to) WQ queues (queue q of n hands out items q, q + n, ... of
// the largest-first order; one counter per queue, so that no address takes an
// atomic per item) or, without queues, a fixed share; it culls both bodies'
// triangles by vertex outcodes into its LDS, makes the kept triangles' float
// boxes there, sweeps their product with each lane on one contiguous range,
// reserves the pair's exact share of the hit arena with one atomic and writes
// the hits in (ta, tb) order.  A pair beyond the wave's limits is put in hlist
// for k_narrow, which writes all of that pair's results: at its front (count at
// queue[NQ_HEAVY]) when it is costly, else at its back (queue[NQ_HEAVY_BACK]).
// hlist holds *P.n entries.
template <int K>
__global__ void __launch_bounds__(NWV * WAVE, 4) k_narrow_wave(DBodies B, DMeshes M, DPairs P, int32_t* hlist,
                                                             unsigned int* queue, int32_t use_queue,
                                                             int32_t heavy_tests, unsigned long long* stamp) {
    stamp_here(stamp);
    ZDX_DYN_LDS(s_nb);
    const int lane = lane_id();
    const int w = threadIdx.x / WAVE;
    char* sl = s_nb + w * WSLICE;
    unsigned char* oc = (unsigned char*)sl;
    const uint64_t below = (1ull << lane) - 1ull;
    const int npairs = *P.n;
    // without queues: the items in rounds of one per wave, each round in the
    // opposite order of the last
    const int nw = (int)gridDim.x * NWV;
    const int gw = (int)blockIdx.x * NWV + w;
    const int nq = nw < WQ ? nw : WQ;   // every queue served by at least one wave
    const int wq = gw % nq;
    auto take = [&](int round) {
        int it = 0;
        if (use_queue) {
            if (lane == 0) it = wq + nq * (int)atomicAdd(queue + wq * QSTRIDE, 1u);
        } else {
            it = round * nw + ((round & 1) ? nw - 1 - gw : gw);
        }
        return it;
    };
#ifdef ZDX_WALK_PROFILE
    const unsigned long long np_entry = wall_clock64();
#endif
    int round = 0;
    int item = uni(__shfl(take(0), 0, WAVE));
    NP_START();
    while (item < npairs) {
        ++round;
        const int next = take(round);   // lane 0's with queues: in flight during the item
        int p, i, j, TA, TB, VA, VB, ndA;
        int64_t oA, oB;
        const int32_t* tiA;
        const int32_t* tiB;
        float boxA[6], boxB[6];
        if (P.items) {
            const NarrowItem& it = P.items[item];
            p = it.p; i = it.i; j = it.j;
            TA = it.TA; TB = it.TB; VA = it.VA; VB = it.VB; ndA = it.nondegA;
            tiA = M.tri_idx + 3 * (int64_t)it.toA;
            tiB = M.tri_idx + 3 * (int64_t)it.toB;
            oA = it.woA; oB = it.woB;
            for (int k = 0; k < 6; ++k) { boxA[k] = it.boxA[k]; boxB[k] = it.boxB[k]; }
        } else {
            p = P.order ? P.order[item] : item;
            i = P.pi[p]; j = P.pj[p];
            const int mA = B.mesh[i], mB = B.mesh[j];
            TA = M.ntris[mA]; TB = M.ntris[mB];
            VA = M.nverts[mA]; VB = M.nverts[mB];
            ndA = B.nondeg[i];
            tiA = M.tri_idx + 3 * (int64_t)M.tri_off[mA];
            tiB = M.tri_idx + 3 * (int64_t)M.tri_off[mB];
            oA = B.wv_off[i]; oB = B.wv_off[j];
            box_f(B.aabb + 6 * (int64_t)i, B.aabb + 6 * (int64_t)i + 3, boxA);
            box_f(B.aabb + 6 * (int64_t)j, B.aabb + 6 * (int64_t)j + 3, boxB);
        }
        const double* wvA = B.wv + 3 * oA;
        const double* wvB = B.wv + 3 * oB;
        NP_MARK(0);
        bool heavy = VA + VB > W_OC_MAX;
        int nA = 0, nB = 0;
        const int ocb = (VA + VB + 15) & ~15;
        const int cap = (WSLICE - ocb) / 40;
        int32_t* ids = (int32_t*)(sl + ocb);   // kept ids: A's, then B's
        if (!heavy) {
            // outcodes of each body's vertices against the other body's box
            constexpr int UV = 4;
            auto vpass = [&](const double* wv, int n, const float* other, unsigned char* ocs) {
                const OcThr th = oc_thresholds(other);
                for (int v0 = 0; v0 < n; v0 += UV * WAVE) {
                    double wq[UV][3];
#pragma unroll
                    for (int u = 0; u < UV; ++u) {
                        const int v = v0 + u * WAVE + lane;
                        if (v < n) { wq[u][0] = wv[3 * v]; wq[u][1] = wv[3 * v + 1]; wq[u][2] = wv[3 * v + 2]; }
                    }
#pragma unroll
                    for (int u = 0; u < UV; ++u) {
                        const int v = v0 + u * WAVE + lane;
                        if (v < n) ocs[v] = (unsigned char)vertex_outcode(wq[u], th);
                    }
                }
            };
            vpass(wvA, VA, boxB, oc);
            vpass(wvB, VB, boxA, oc + VA);
            wave_sync();
            NP_MARK(1);
            // the kept triangles of each side, in triangle order, four layers
            // of 64 triangles at a time (their index loads in flight together)
            auto cull = [&](const int32_t* ti, int T, const unsigned char* ocs, int base, int& nk) {
                for (int t0 = 0; t0 < T; t0 += 4 * WAVE) {
                    int ix[4][3];
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const int t = t0 + u * WAVE + lane;
                        if (t < T) { const Tri3 q = ld_tri(ti, t); ix[u][0] = q.a; ix[u][1] = q.b; ix[u][2] = q.c; }
                    }
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const int t = t0 + u * WAVE + lane;
                        const bool keep = t < T && (ocs[ix[u][0]] & ocs[ix[u][1]] & ocs[ix[u][2]]) == 0;
                        const uint64_t m = __ballot(keep);
                        const int r = base + nk + __popcll(m & below);
                        if (keep && r < cap) ids[r] = t;
                        nk += __popcll(m);
                    }
                }
            };
            cull(tiA, TA, oc, 0, nA);
            cull(tiB, TB, oc + VA, nA, nB);
            heavy = nA + nB > cap || (int64_t)nA * nB > heavy_tests;
            NP_MARK(2);
        }
        if (heavy) {
            if (lane == 0) {
                if (VA + VB > W_OC_MAX || (int64_t)nA * nB >= W_HEAVY_FIRST) hlist[atomicAdd(queue + NQ_HEAVY, 1u)] = item;
                else hlist[npairs - 1 - (int)atomicAdd(queue + NQ_HEAVY_BACK, 1u)] = item;
            }
        } else {
            wave_sync();
            // float boxes and world-vertex indices of the kept triangles
            const int nk = nA + nB;
            float* bx = (float*)(ids + nk);       // A's boxes, then B's
            int32_t* v3 = (int32_t*)(bx + 6 * nk);
            for (int k = lane; k < nk; k += WAVE) {
                const bool a = k < nA;
                const int t = ids[k];
                const int32_t* ix = (a ? tiA : tiB) + 3 * t;
                const int i0 = ix[0], i1 = ix[1], i2 = ix[2];
                const double* wv = a ? wvA : wvB;
                V3 tr[3] = {ld3(wv + 3 * i0), ld3(wv + 3 * i1), ld3(wv + 3 * i2)};
                double lo[3], hi[3];
                tri_box(tr, lo, hi);
                box_f(lo, hi, bx + 6 * k);
                const int64_t o = a ? oA : oB;
                v3[3 * k] = (int32_t)(o + i0); v3[3 * k + 1] = (int32_t)(o + i1); v3[3 * k + 2] = (int32_t)(o + i2);
            }
            wave_sync();
            NP_MARK(3);
            // ordered sweep: each lane one contiguous range of the product, a
            // counting pass (which keeps the hits as bits when the range has
            // at most 64 tests), a wave scan, one reservation, the write pass
            // (from the bits, or by testing again)
            const float* FA = bx;
            const float* FB = bx + 6 * nA;
            const int total = nA * nB;
            const int per = (total + WAVE - 1) / WAVE;
            const int g0 = lane * per < total ? lane * per : total;
            const int g1 = g0 + per < total ? g0 + per : total;
            int cnt = 0;
            uint64_t bits = 0ull;
            sweep_range<K>(g0, g1 - g0, nB, FA, FB, [&](int k, int, int) {
                ++cnt;
                if (k < 64) bits |= 1ull << k;
            });
            int incl = cnt;
#pragma unroll
            for (int d = 1; d < WAVE; d <<= 1) {
                const int y = __shfl_up(incl, d, WAVE);
                if (lane >= d) incl += y;
            }
            const int hits = uni(__shfl(incl, WAVE - 1, WAVE));
            unsigned long long hb = 0ull;
            if (lane == 0) hb = reserve_hits(P, item, hits);
            hb = __shfl(hb, 0, WAVE);
            NP_MARK(4);
            const bool ok = hb != ~0ull;
            if (ok && cnt > 0) {
                int64_t at = (int64_t)hb + (incl - cnt);
                // the hit records: the world-vertex indices of (ta, tb)
                auto emit = [&](int ca, int cb) {
                    const int32_t* va = v3 + 3 * ca;
                    const int32_t* vb = v3 + 3 * (nA + cb);
                    int2* dst = (int2*)(P.hits + 6 * at);
                    dst[0] = make_int2(va[0], va[1]);
                    dst[1] = make_int2(va[2], vb[0]);
                    dst[2] = make_int2(vb[1], vb[2]);
                    if (P.hit_tri) ((int2*)P.hit_tri)[at] = make_int2(ids[ca], ids[nA + cb]);
                    ++at;
                };
                if (per <= 64) {
                    for (uint64_t m = bits; m; m &= m - 1) {
                        const int g = g0 + ctz64(m);
                        const int ca = g / nB;
                        emit(ca, g - ca * nB);
                    }
                } else {
                    sweep_range<K>(g0, g1 - g0, nB, FA, FB, [&](int, int ca, int cb) { emit(ca, cb); });
                }
            }
            NP_MARK(5);
            if (lane == 0) {
                P.hit_off[p] = ok ? (int64_t)hb : 0;
                P.tri_hits[p] = ok ? hits : 0;
                P.tri_checks[p] = (int64_t)ndA * TB;
                P.tri_tests[p] = total;
                P.tri_scanned[p] = TA + TB;
                P.seg_off[p] = 0;   // set by the compaction for pairs with hits
                P.nseg[p] = 0;
            }
        }
        // the wave's LDS is rewritten by the next item
        wave_sync();
        NP_MARK(6);
#ifdef ZDX_WALK_PROFILE
        if (lane == 0 && item < 16384) {
            long long* r = g_np_pairs + NP_REC * (int64_t)item;
            r[0] = nA; r[1] = nB; r[2] = heavy ? -1 : 0; r[3] = p;
            r[4] = TA; r[5] = TB; r[6] = VA; r[7] = VB;
            long long tot = 0;
            for (int k = 0; k < 8; ++k) { r[8 + k] = np_ph[k]; tot += np_ph[k]; np_ph[k] = 0; }
            r[16] = gw; r[17] = (long long)np_t0 - tot; r[18] = (long long)np_t0; r[19] = 1 + 2 * (long long)__smid();
            if (round == 1) r[2] = (long long)np_entry;   // the wave's first instruction (first item only)
        }
#endif
        item = uni(__shfl(next, 0, WAVE));
    }
}
