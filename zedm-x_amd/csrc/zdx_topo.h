// zdx_topo.h — segment topology of one contact pair, split into phases.
//
// The reference (src/main.cpp:2152-2219) runs, per pair: components
// (label_segments_by_contact_greedy 804-898, when split), then per component
// snap_endpoints (680-746), split_t_junctions (900-961), extract_loops
// (963-1025), accumulate_Sn_Gn_from_polyline (1027-1040), contact_point_xc0
// (1042-1052), the <3-vertex and area_eps filters and the flip of nA toward B
// (2188-2219).
//
// Here the work is cut where the quadratic searches are: every test that
// does not depend on another (segment pair adjacency, endpoint pair snap
// distance, (edge, vertex) T-junction test) is a pure function of one item,
// so the GPU runs all items of all pairs as flat passes over the device, and
// only the order-dependent walks run serially per pair:
//   pass A  (items)  tp_cells, tp_adj_key, tp_snap_hit
//   walk B  (pair)   tp_components_and_vertices: DFS components, union-find
//                    over the snap hits of each component, first-occurrence
//                    vertex ids, sequential vertex means, edges
//   pass C  (items)  tp_tj_hit
//   walk D  (pair)   tp_loops: T-junction chains, loop tracing, Sn/Gn/xc0,
//                    filters, sink
// tp_topology_host() runs the same functions in sequence on the host (tests).
//
// Walks B and D execute "uniformly redundantly": every lane of a wave runs
// the same serial code on identical values; WP::ballot spreads a scan of up
// to 64 items over the lanes.  With LaneHost (W = 1) the same code is the
// plain serial algorithm.
#pragma once
#include <vector>

#include "zdx_geom.h"

namespace zdx {

constexpr int TP_NOADJ = 255;   // adjacency byte of a pair that is not joined

// ---------------------------------------------------------------- pass A
// Cells of a segment's representative points p0, p1, mid (cell = tol).
ZDX_HD inline void tp_cells(const double* seg, int s, double h, int* c9) {
    const V3 p0 = v3(seg[6 * s + 0], seg[6 * s + 1], seg[6 * s + 2]);
    const V3 p1 = v3(seg[6 * s + 3], seg[6 * s + 4], seg[6 * s + 5]);
    const V3 mid = mul(add(p0, p1), 0.5);
    const I3 a = cell_of(p0, h), b = cell_of(p1, h), c = cell_of(mid, h);
    c9[0] = a.x; c9[1] = a.y; c9[2] = a.z;
    c9[3] = b.x; c9[4] = b.y; c9[5] = b.z;
    c9[6] = c.x; c9[7] = c.y; c9[8] = c.z;
}

// Component candidate test of segment j found from popped segment i
// (main.cpp:834-879): j in the +-1 cell neighbourhood of one of i's
// representative points, AABB(+tol) overlap, seg_seg_dist2 <= tol^2.
// Returns the reference's scan-order key (rep point r of i, neighbour cell
// (dx,dy,dz) lexicographic: r*27 + cell offset, 0..80) or TP_NOADJ.
ZDX_HD inline int tp_adj_key(const double* seg, const int* cell9, double h, int i, int j) {
    if (i == j) return TP_NOADJ;
    int key = TP_NOADJ;
    for (int r = 0; r < 3; ++r) {
        const int ax = cell9[9 * i + 3 * r], ay = cell9[9 * i + 3 * r + 1], az = cell9[9 * i + 3 * r + 2];
        for (int q = 0; q < 3; ++q) {
            const int bx = cell9[9 * j + 3 * q], by = cell9[9 * j + 3 * q + 1], bz = cell9[9 * j + 3 * q + 2];
            if (near1(ax, ay, az, bx, by, bz)) {
                const int k = r * 27 + (bx - ax + 1) * 9 + (by - ay + 1) * 3 + (bz - az + 1);
                if (k < key) key = k;
            }
        }
    }
    if (key == TP_NOADJ) return TP_NOADJ;
    const V3 p1 = v3(seg[6 * i + 0], seg[6 * i + 1], seg[6 * i + 2]), q1 = v3(seg[6 * i + 3], seg[6 * i + 4], seg[6 * i + 5]);
    const V3 p2 = v3(seg[6 * j + 0], seg[6 * j + 1], seg[6 * j + 2]), q2 = v3(seg[6 * j + 3], seg[6 * j + 4], seg[6 * j + 5]);
    const V3 amn = v3(dmin(p1.x, q1.x) - h, dmin(p1.y, q1.y) - h, dmin(p1.z, q1.z) - h);
    const V3 amx = v3(dmax(p1.x, q1.x) + h, dmax(p1.y, q1.y) + h, dmax(p1.z, q1.z) + h);
    const V3 bmn = v3(dmin(p2.x, q2.x) - h, dmin(p2.y, q2.y) - h, dmin(p2.z, q2.z) - h);
    const V3 bmx = v3(dmax(p2.x, q2.x) + h, dmax(p2.y, q2.y) + h, dmax(p2.z, q2.z) + h);
    if (amx.x < bmn.x || amn.x > bmx.x || amx.y < bmn.y || amn.y > bmx.y || amx.z < bmn.z || amn.z > bmx.z)
        return TP_NOADJ;
    return seg_seg(p1, q1, p2, q2) <= h * h ? key : TP_NOADJ;
}

// Row i of the adjacency as a list: the accepted candidates j of popped
// segment i in ascending j, each as j | key << 24; writes at most R entries
// and returns the true count (a count above R means the list did not fit).
ZDX_HD inline int tp_adj_row(const double* seg, const int* cell9, double h, int S, int i, int* out, int R) {
    int m = 0;
    for (int j = 0; j < S; ++j) {
        const int key = tp_adj_key(seg, cell9, h, i, j);
        if (key == TP_NOADJ) continue;
        if (m < R) out[m] = j | (key << 24);
        ++m;
    }
    return m;
}

// Endpoint pairs are numbered t = a(a-1)/2 + b, b < a; point k is end k&1 of
// segment k>>1.  The snap test of snap_endpoints (main.cpp:697-723): cells
// within +-1 and squared distance <= tol^2.
ZDX_HD inline void tp_pair_of(int64_t t, int& a, int& b) {
    int i = (int)((1.0 + sqrt(1.0 + 8.0 * (double)t)) * 0.5);
    while ((int64_t)i * (i - 1) / 2 > t) --i;
    while ((int64_t)(i + 1) * i / 2 <= t) ++i;
    a = i;
    b = (int)(t - (int64_t)i * (i - 1) / 2);
}
ZDX_HD inline bool tp_snap_hit(const double* seg, const int* cell9, double h, int a, int b) {
    const int* ca = cell9 + 9 * (a >> 1) + 3 * (a & 1);
    const int* cb = cell9 + 9 * (b >> 1) + 3 * (b & 1);
    if (!near1(ca[0], ca[1], ca[2], cb[0], cb[1], cb[2])) return false;
    const double* pa = seg + 6 * (a >> 1) + 3 * (a & 1);
    const double* pb = seg + 6 * (b >> 1) + 3 * (b & 1);
    return norm2(sub(v3(pa[0], pa[1], pa[2]), v3(pb[0], pb[1], pb[2]))) <= h * h;
}

// ---------------------------------------------------------------- walk B
// Per processed component q (in the reference's processing order: size
// descending, stable), 5 ints: vertex count, edge count, first vertex, first
// edge (both within the pair), first T-junction bit word (within the pair).
enum { TPC_NV = 0, TPC_NE = 1, TPC_VB = 2, TPC_EB = 3, TPC_TW = 4, TPC_N = 5 };

// Scratch of walk B for a pair of S segments and adjacency rows of R entries
// (the rows themselves live in the caller's buffer): ints.
ZDX_HD inline int64_t tp_b_ints(int S) { return 24 * (int64_t)S + 4; }
// T-junction words of a component: one row of ceil(nv/64) words per edge.
ZDX_HD inline int tp_row_words(int nv) { return (nv + 63) >> 6; }

// adjn/adjl: per segment the count and the list (R entries, j | key << 24,
// ascending j) of its accepted component candidates; snapw: bit t of word
// t/64 is the snap hit of endpoint pair t.  Outputs: comp (ncomp x TPC_N), V
// (3 per vertex), vcell (3 per vertex), E (2 per edge, vertex ids local to the
// component).  Returns ncomp; *tj_words = total T-junction words.
// Optional sensitivities: EJ (36 per endpoint, point order 2s+e) gives VJ (36
// per vertex, same numbering as V): the mean of its endpoints' Jacobians,
// summed in point order.
template <class WP>
ZDX_HD inline int tp_components_and_vertices(const WP& wp, const double* seg, int S, double h, bool split,
                                             const int* adjn, const int* adjl, int R, const uint64_t* snapw, int* I,
                                             int* comp, double* V, int* vcell, int* E, int* tj_words,
                                             const double* EJ = nullptr, double* VJ = nullptr,
                                             const int* glab = nullptr) {
    const int W = WP::W;
    const int ln = wp.lane();
    const int n = 2 * S;
    int* vis = I;            // S
    int* stk = vis + S;      // S
    int* spk = stk + S;      // S
    int* cord = spk + S;     // S   segment order of all comps, concatenated
    int* cbeg = cord + S;    // S+1
    int* csz = cbeg + S + 1; // S+1
    int* cidx = csz + S + 1; // S   comp processing order
    int* compof = cidx + S;  // S   comp of each segment
    int* par = compof + S;   // 2S
    int* rnk = par + n;      // 2S
    int* rootvid = rnk + n;  // 2S
    int* ptv = rootvid + n;  // 2S
    int* cnt = ptv + n;      // 2S
    int* voff = cnt + n;     // 2S+1  members of vertex v: mem[voff[v] .. voff[v+1])
    int* mem = voff + n + 1; // 2S    points of the component, grouped by vertex, in point order
    int* vfill = mem + n;    // 2S
    // ---- components (DFS in the reference's candidate order)
    int ncomp = 0;
    if (split) {
        for (int s = ln; s < S; s += W) vis[s] = 0;
        wp.sync();
        int nord = 0;
        for (int seed = 0; seed < S; ++seed) {
            if (vis[seed]) continue;
            cbeg[ncomp] = nord;
            int sp = 0;
            stk[sp++] = seed;
            vis[seed] = 1;
            cord[nord++] = seed;
            int top = seed;   // stk[sp - 1], held in a register after a push
            while (sp > 0) {
                const int i = top;
                --sp;
                const int first = sp;
                const int* row = adjl + (int64_t)i * R;
                const int m = adjn[i];
                const int e0 = ln < R ? row[ln] : 0;   // loaded with the count
                if (m <= W) {
                    // lanes take the row's entries (ascending j); the accepted,
                    // unvisited ones go on the stack ordered by (scan key, id)
                    const int e = ln < m ? e0 : 0;
                    const int jj = e & 0xFFFFFF, kj = e >> 24;
                    const bool acc = ln < m && !vis[jj];
                    const uint64_t am = wp.ballot(acc);
                    int rank = 0;
                    for (uint64_t b = am; b; b &= b - 1) {
                        const int l = ctz64(b);
                        const int kl = wp.readlane(kj, l);
                        rank += (kl < kj || (kl == kj && l < ln)) ? 1 : 0;
                    }
                    const int na = __builtin_popcountll(am);
                    wp.sync();
                    if (acc) {
                        stk[first + rank] = jj;
                        vis[jj] = 1;
                        cord[nord + rank] = jj;
                    }
                    // the new top: the accepted lane of the highest rank
                    int tl = 0;
                    if (na > 0) {
                        const uint64_t hi = wp.ballot(acc && rank == na - 1);
                        tl = wp.readlane(jj, ctz64(hi));
                    }
                    wp.sync();
                    sp += na;
                    nord += na;
                    top = na > 0 ? tl : (sp > 0 ? stk[sp - 1] : 0);
                    continue;
                }
                for (int q = 0; q < m; ++q) {
                    const int e = row[q];
                    const int jj = e & 0xFFFFFF, kj = e >> 24;
                    if (vis[jj]) continue;
                    // accepted candidates of i are pushed ordered by (scan key, id)
                    int pos = sp;
                    while (pos > first && spk[pos - 1] > kj) {
                        stk[pos] = stk[pos - 1];
                        spk[pos] = spk[pos - 1];
                        --pos;
                    }
                    stk[pos] = jj;
                    spk[pos] = kj;
                    ++sp;
                }
                for (int k = first; k < sp; ++k) {
                    vis[stk[k]] = 1;
                    cord[nord++] = stk[k];
                }
                top = sp > 0 ? stk[sp - 1] : 0;
            }
            csz[ncomp] = nord - cbeg[ncomp];
            ++ncomp;
        }
        for (int c = 0; c < ncomp; ++c) cidx[c] = c;
        CompBySizeDesc cmp; cmp.size = SArr<int>{csz, 1};
        ref_sort(SArr<int>{cidx, 1}, ncomp, cmp);
        for (int c = 0; c < ncomp; ++c)
            for (int k = ln; k < csz[c]; k += W) compof[cord[cbeg[c] + k]] = c;
    } else {
        for (int s = ln; s < S; s += W) { cord[s] = s; compof[s] = 0; }
        cbeg[0] = 0; csz[0] = S; cidx[0] = 0;
        ncomp = 1;
    }
    wp.mark(1);
    // ---- snap: connected components of the snap hits inside each component
    // (the reference's union-find partition, which does not depend on the
    // order of the unions).  Lanes take the hit words in parallel and link
    // roots lock-free: a root only ever points to a smaller index (compare and
    // swap), so lab[] stays a forest; then every point is pointed at its root.
    int* lab = par;            // 2S
    // glab: the snap clusters of the whole pair found beforehand (each point's
    // smallest point), used when every cluster lies in one component: then
    // no snap hit joins two components and they are the same partition
    bool pre = false;
    if (glab != nullptr && glab[0] >= 0) {
        for (int k = ln; k < n; k += W) lab[k] = glab[k];
        wp.sync();
        bool bad = false;
        if (split && ncomp > 1)
            for (int k = ln; k < n; k += W) bad = bad || compof[k >> 1] != compof[lab[k] >> 1];
        pre = wp.ballot(bad) == 0ull;
    }
    // find with pointer jumping: every node passed is pointed at its
    // grandparent.  A parent always has the smaller index and a node that is
    // not a root never becomes one again, so these plain stores never touch a
    // root (which only the compare and swap changes) and only ever point a
    // node at one of its ancestors: the forest and its partition are kept.
    auto root = [&](int x) {
        int y = lab[x];
        if (y == x) return x;
        int z = lab[y];
        while (z != y) {
            lab[x] = z;
            x = y; y = z; z = lab[y];
        }
        return y;
    };
    if (!pre) {
        for (int k = ln; k < n; k += W) lab[k] = k;
        wp.sync();
        wp.mark(5);
        const int64_t npairs = (int64_t)n * (n - 1) / 2;
        const int64_t nwords = (npairs + 63) >> 6;
        // the set bits of each round of W words are spread over the lanes through
        // a list of up to W hit indices (in rnk.., free until the numbering below:
        // min(W, hits) <= n (n - 1) / 2 and <= 7 n ints are there), so the unions
        // run on all lanes at once instead of one lane per word
        int* hl = rnk;
        auto unite = [&](int64_t t) {
            int a, b;
            tp_pair_of(t, a, b);
            if (compof[a >> 1] != compof[b >> 1]) return;
            int ra = root(a), rb = root(b);
            while (ra != rb) {
                if (ra < rb) { const int x = ra; ra = rb; rb = x; }   // link the larger root under the smaller
                const int old = wp.atomic_cas(&lab[ra], ra, rb);
                if (old == ra) break;
                ra = root(old);
                rb = root(rb);
            }
        };
        // the words of SNAP_BATCH rounds are loaded together, the next batch's
        // loads in flight while this batch's bits are united
        constexpr int SNAP_BATCH = 4;
        uint64_t nxt[SNAP_BATCH];
        for (int r = 0; r < SNAP_BATCH; ++r) {
            const int64_t w = (int64_t)r * W + ln;
            nxt[r] = w < nwords ? snapw[w] : 0ull;
        }
        for (int64_t wb0 = 0; wb0 < nwords; wb0 += (int64_t)SNAP_BATCH * W) {
            uint64_t cur[SNAP_BATCH];
            for (int r = 0; r < SNAP_BATCH; ++r) {
                cur[r] = nxt[r];
                const int64_t w = wb0 + (int64_t)(SNAP_BATCH + r) * W + ln;
                nxt[r] = w < nwords ? snapw[w] : 0ull;
            }
            wp.mark(4);
            for (int r = 0; r < SNAP_BATCH; ++r) {
                const int64_t wb = wb0 + (int64_t)r * W;
                if (wb >= nwords) break;
                const int64_t w = wb + ln;
                uint64_t word = cur[r];
                if (W == 1) {   // one lane: take the bits in order
                    while (word) {
                        const int64_t t = (w << 6) + ctz64(word);
                        word &= word - 1;
                        unite(t);
                    }
                    continue;
                }
                int nbits;
                int g = wp.scan_excl(__builtin_popcountll(word), nbits);
                for (int j0 = 0; j0 < nbits; j0 += W) {
                    while (word && g < j0 + W) {
                        hl[g - j0] = (int)(((w - wb) << 6) + ctz64(word));   // index within the round
                        word &= word - 1;
                        ++g;
                    }
                    wp.sync();
                    if (j0 + ln < nbits) unite((wb << 6) + hl[ln]);
                    wp.sync();
                }
            }
        }
        wp.mark(6);
    }
    wp.sync();
    wp.mark(7);
    for (int k = ln; k < n; k += W) lab[k] = root(k);
    for (int k = ln; k < n; k += W) rootvid[k] = -1;
    wp.sync();
    auto find = [&](int x) { return lab[x]; };
    wp.mark(2);
    // ---- per component: first-occurrence vertex ids, then each vertex's
    // mean over its points summed in point order (one lane per vertex), edges
    int vb = 0, eb = 0, tw = 0;
    for (int q = 0; q < ncomp; ++q) {
        const int c = cidx[q];
        const int base = cbeg[c], ns = csz[c];
        const int np = 2 * ns;
        auto ptk = [&](int k) { return 2 * cord[base + (k >> 1)] + (k & 1); };
        // first position of every cluster in the component's point order
        int* firstp = rnk;
        for (int k = ln; k < np; k += W) firstp[find(ptk(k))] = np;
        wp.sync();
        for (int k = ln; k < np; k += W) wp.atomic_min(&firstp[find(ptk(k))], k);
        wp.sync();
        wp.mark(8);
        // vertex id of a cluster = number of clusters that start before it
        int nv = 0;
        for (int kb = 0; kb < np; kb += W) {
            const int k = kb + ln;
            const int r = k < np ? find(ptk(k)) : 0;
            const bool f = k < np && firstp[r] == k;
            const uint64_t fm = wp.ballot(f);
            if (f) rootvid[r] = nv + (int)__builtin_popcountll(fm & ((1ull << ln) - 1ull));
            nv += (int)__builtin_popcountll(fm);
        }
        wp.mark(9);
        for (int v = ln; v < nv; v += W) cnt[v] = 0;
        wp.sync();
        for (int k = ln; k < np; k += W) {
            const int pt = ptk(k);
            const int v = rootvid[find(pt)];
            ptv[pt] = v;
            wp.atomic_add(&cnt[v], 1);
        }
        wp.sync();
        wp.mark(10);
        int acc = 0;
        for (int vb0 = 0; vb0 < nv; vb0 += W) {
            const int v = vb0 + ln;
            const int c = v < nv ? cnt[v] : 0;
            int tot;
            const int off = wp.scan_excl(c, tot);
            if (v < nv) { voff[v] = acc + off; vfill[v] = acc + off; }
            acc += tot;
        }
        if (ln == 0) voff[nv] = acc;
        wp.sync();
        wp.mark(11);
        // points grouped by vertex in point order: per chunk of W points, each
        // point's rank among the earlier points of its vertex in the chunk
        // (lanes with the same vertex found by ballots over the bits of the
        // vertex id), then the lowest such lane advances the vertex's cursor
        for (int kb = 0; kb < np; kb += W) {
            const int k = kb + ln;
            const bool act = k < np;
            const int pt = act ? ptk(k) : 0;
            const int v = act ? ptv[pt] : 0;
            uint64_t same = wp.ballot(act);
            for (int bit = 0; (1 << bit) < nv; ++bit) {
                const bool on = ((v >> bit) & 1) != 0;
                const uint64_t bm = wp.ballot(act && on);
                same &= on ? bm : ~bm;
            }
            const uint64_t below = same & ((1ull << ln) - 1ull);
            if (act) mem[vfill[v] + (int)__builtin_popcountll(below)] = pt;
            wp.sync();
            if (act && below == 0ull) vfill[v] += (int)__builtin_popcountll(same);
            wp.sync();
        }
        wp.mark(12);
        for (int v = ln; v < nv; v += W) {
            const int m0 = voff[v], m1 = voff[v + 1];
            const double* p = seg + 3 * mem[m0];     // point pt at seg + 6 (pt >> 1) + 3 (pt & 1) = seg + 3 pt
            double sx = p[0], sy = p[1], sz = p[2];
            int m = m0 + 1;
            for (; m + 4 <= m1; m += 4) {   // four members' loads in flight, summed in member order
                const double* q0 = seg + 3 * mem[m];
                const double* q1 = seg + 3 * mem[m + 1];
                const double* q2 = seg + 3 * mem[m + 2];
                const double* q3 = seg + 3 * mem[m + 3];
                const double x0 = q0[0], y0 = q0[1], z0 = q0[2], x1 = q1[0], y1 = q1[1], z1 = q1[2];
                const double x2 = q2[0], y2 = q2[1], z2 = q2[2], x3 = q3[0], y3 = q3[1], z3 = q3[2];
                sx += x0; sy += y0; sz += z0;
                sx += x1; sy += y1; sz += z1;
                sx += x2; sy += y2; sz += z2;
                sx += x3; sy += y3; sz += z3;
            }
            for (; m < m1; ++m) {
                const double* pm = seg + 3 * mem[m];
                sx += pm[0]; sy += pm[1]; sz += pm[2];
            }
            const double cn = (double)(m1 - m0);
            const V3 mv = v3(sx / cn, sy / cn, sz / cn);
            double* o = V + 3 * (int64_t)(vb + v);
            o[0] = mv.x; o[1] = mv.y; o[2] = mv.z;
            const I3 cc = cell_of(mv, h);
            int* oc = vcell + 3 * (int64_t)(vb + v);
            oc[0] = cc.x; oc[1] = cc.y; oc[2] = cc.z;
        }
        if (VJ)
            for (int v = 0; v < nv; ++v)
                for (int cj = ln; cj < 36; cj += W) {
                    double a = EJ[36 * (int64_t)mem[voff[v]] + cj];
                    for (int m = voff[v] + 1; m < voff[v + 1]; ++m) a += EJ[36 * (int64_t)mem[m] + cj];
                    VJ[36 * (int64_t)(vb + v) + cj] = a / (double)(voff[v + 1] - voff[v]);
                }
        wp.mark(13);
        int ne = 0;
        for (int kb = 0; kb < ns; kb += W) {   // edges in segment order (ballot prefix)
            const int k = kb + ln;
            int a = 0, b = 0;
            if (k < ns) {
                const int s = cord[base + k];
                a = ptv[2 * s];
                b = ptv[2 * s + 1];
            }
            const bool e = k < ns && a != b;
            const uint64_t m = wp.ballot(e);
            if (e) {
                const int r = eb + ne + (int)__builtin_popcountll(m & ((1ull << ln) - 1ull));
                E[2 * (int64_t)r] = a;
                E[2 * (int64_t)r + 1] = b;
            }
            ne += (int)__builtin_popcountll(m);
        }
        if (ln == 0) {
            int* cr = comp + TPC_N * q;
            cr[TPC_NV] = nv; cr[TPC_NE] = ne; cr[TPC_VB] = vb; cr[TPC_EB] = eb; cr[TPC_TW] = tw;
        }
        vb += nv;
        eb += ne;
        tw += ne * tp_row_words(nv);
        wp.sync();
    }
    wp.mark(3);
    *tj_words = tw;
    return ncomp;
}

// ---------------------------------------------------------------- pass C
// split_t_junctions candidate test (main.cpp:908-935) of vertex v against
// edge e of one component (V, vcell: that component's vertices), in two
// parts: the cell range of the edge's box grown by h (cb: 6 ints, lower then
// upper cell), which depends on the edge only, and the test of one vertex.
ZDX_HD inline void tp_tj_box(const double* V, const int* E, double h, int e, int* cb) {
    const int a = E[2 * e], b = E[2 * e + 1];
    const V3 pa = v3(V[3 * a], V[3 * a + 1], V[3 * a + 2]), pb = v3(V[3 * b], V[3 * b + 1], V[3 * b + 2]);
    const V3 mn = v3(dmin(pa.x, pb.x) - h, dmin(pa.y, pb.y) - h, dmin(pa.z, pb.z) - h);
    const V3 mx = v3(dmax(pa.x, pb.x) + h, dmax(pa.y, pb.y) + h, dmax(pa.z, pb.z) + h);
    const I3 c0 = cell_of(mn, h), c1 = cell_of(mx, h);
    cb[0] = c0.x; cb[1] = c0.y; cb[2] = c0.z; cb[3] = c1.x; cb[4] = c1.y; cb[5] = c1.z;
}
// the distance test of vertex v against the edge (a, b), after the cell test
ZDX_HD inline bool tp_tj_near(const double* V, double h, int a, int b, int v) {
    const V3 pa = v3(V[3 * a], V[3 * a + 1], V[3 * a + 2]), pb = v3(V[3 * b], V[3 * b + 1], V[3 * b + 2]);
    double d2, t;
    pt_seg(v3(V[3 * v], V[3 * v + 1], V[3 * v + 2]), pa, pb, d2, t);
    return d2 <= h * h && t > 1e-6 && t < 1.0 - 1e-6;
}
ZDX_HD inline bool tp_tj_test(const double* V, const int* vcell, const int* E, double h, int e, int v, const int* cb) {
    const int a = E[2 * e], b = E[2 * e + 1];
    if (v == a || v == b) return false;
    const int cx = vcell[3 * v], cy = vcell[3 * v + 1], cz = vcell[3 * v + 2];
    if (cx < cb[0] || cx > cb[3] || cy < cb[1] || cy > cb[4] || cz < cb[2] || cz > cb[5]) return false;
    return tp_tj_near(V, h, a, b, v);
}
ZDX_HD inline bool tp_tj_hit(const double* V, const int* vcell, const int* E, double h, int e, int v) {
    int cb[6];
    tp_tj_box(V, E, h, e, cb);
    return tp_tj_test(V, vcell, E, h, e, v, cb);
}

// ---------------------------------------------------------------- walk D
// Scratch of walk D for a pair of S segments (vertices <= 2S, split edges <= 4S).
ZDX_HD inline int64_t tp_d_ints(int S) { return 4 * (int64_t)(2 * S) + 2 * (4 * (int64_t)S) * 3 + 2 * (4 * (int64_t)S) + 8; }
ZDX_HD inline int64_t tp_d_dbls(int S) { return 3 * (2 * (int64_t)S) + 4 * (int64_t)S + 8; }

// Sensitivities of a loop's Sn and Gn (accumulate_Sn_Gn_from_polyline) to
// the motion parameters, from the vertex Jacobians VJ (36 per vertex):
//   dSn = -1/2 sum_m [p_{m+1} - p_{m-1}]x J_m
//   dGn = sum_m dg/dp0 J_m + dg/dp1 J_{m+1},  g = -1/3 dx (p0.p1 + 1/3 |dx|^2)
// out: 72 doubles (dSn 3 x 12, then dGn 3 x 12); entry c computed by lane c mod W.
template <class WP>
ZDX_HD inline void tp_loop_jacobian(const WP& wp, const int* path, int plen, const double* Vs, const double* VJ,
                                    double* out) {
    const int W = WP::W;
    for (int c = wp.lane(); c < 72; c += W) {
        const int row = (c % 36) / 12, col = c % 12;
        double acc = 0.0;
        for (int m = 0; m < plen; ++m) {
            const int vm = path[m], vn = path[(m + 1) % plen], vp = path[(m + plen - 1) % plen];
            const double* Jm = VJ + 36 * (int64_t)vm;
            if (c < 36) {
                // -1/2 [w]x J,  w = p_{m+1} - p_{m-1}
                const double w0 = Vs[3 * vn] - Vs[3 * vp], w1 = Vs[3 * vn + 1] - Vs[3 * vp + 1],
                             w2 = Vs[3 * vn + 2] - Vs[3 * vp + 2];
                const double j0 = Jm[col], j1 = Jm[12 + col], j2 = Jm[24 + col];
                const double cr = row == 0 ? (w1 * j2 - w2 * j1) : (row == 1 ? (w2 * j0 - w0 * j2) : (w0 * j1 - w1 * j0));
                acc += -0.5 * cr;
            } else {
                const double* Jn = VJ + 36 * (int64_t)vn;
                const V3 p0 = v3(Vs[3 * vm], Vs[3 * vm + 1], Vs[3 * vm + 2]);
                const V3 p1 = v3(Vs[3 * vn], Vs[3 * vn + 1], Vs[3 * vn + 2]);
                const V3 dx = sub(p1, p0);
                const double sv = dot(p0, p1) + (1.0 / 3.0) * dot(dx, dx);
                const V3 a0 = sub(p1, mul(dx, 2.0 / 3.0));   // ds/dp0
                const V3 a1 = add(p0, mul(dx, 2.0 / 3.0));   // ds/dp1
                const double dxr = row == 0 ? dx.x : (row == 1 ? dx.y : dx.z);
                // (dg/dp0 J_m)[row][col] = 1/3 s J_m[row] - 1/3 dx_row (a0 . J_m[:,col])
                const double a0J = a0.x * Jm[col] + a0.y * Jm[12 + col] + a0.z * Jm[24 + col];
                const double a1J = a1.x * Jn[col] + a1.y * Jn[12 + col] + a1.z * Jn[24 + col];
                acc += (1.0 / 3.0) * sv * Jm[12 * row + col] - (1.0 / 3.0) * dxr * a0J;
                acc += -(1.0 / 3.0) * sv * Jn[12 * row + col] - (1.0 / 3.0) * dxr * a1J;
            }
        }
        out[c] = acc;
    }
}

// comp/V/vcell/E as written by walk B, tjw: the T-junction bit words of the
// pair.  For every accepted loop, sink(loop_index, comp, raw_index, Sn, Gn,
// nA, xc0, area, nverts, path, Vs, jac) is called (by every lane) in the
// reference's order; jac (72 doubles, tp_loop_jacobian) when VJ (vertex
// Jacobians of the pair, walk B) is given, else nullptr.
template <class WP, class Sink>
ZDX_HD inline TopoOut tp_loops(const WP& wp, int S, int ncomp, double h, double area_eps, const V3& dirAB,
                               const int* comp, const double* V, const int* vcell, const int* E, const uint64_t* tjw,
                               int* I, double* D, Sink& sink, const double* VJ = nullptr, double* jac = nullptr) {
    TopoOut out; out.comps = ncomp; out.loops_raw = 0; out.loops = 0; out.status = TOPO_OK;
    const int W = WP::W;
    const int ln = wp.lane();
    const int NV = 2 * S, E2c = 4 * S;
    int* deg = I;                // NV
    int* aoff = deg + NV;        // NV
    int* afill = aoff + NV;      // NV
    int* sord = afill + NV;      // NV
    int* E2 = sord + NV;         // 2 E2c
    int* adjn = E2 + 2 * E2c;    // 2 E2c
    int* adje = adjn + 2 * E2c;  // 2 E2c
    int* used = adje + 2 * E2c;  // E2c
    int* path = used + E2c;      // E2c + 1
    double* Vs = D;              // 3 NV  this component's vertices
    double* ht = Vs + 3 * NV;    // E2c  T-junction hits: parameter t
    for (int q = 0; q < ncomp; ++q) {
        const int* cr = comp + TPC_N * q;
        const int nv = cr[TPC_NV], ne = cr[TPC_NE];
        const double* Vq = V + 3 * (int64_t)cr[TPC_VB];
        const int* vcq = vcell + 3 * (int64_t)cr[TPC_VB];
        const int* Eq = E + 2 * (int64_t)cr[TPC_EB];
        const uint64_t* Tq = tjw + cr[TPC_TW];
        for (int k = ln; k < 3 * nv; k += W) Vs[k] = Vq[k];
        wp.sync();
        auto VV = [&](int v) { return v3(Vs[3 * v], Vs[3 * v + 1], Vs[3 * v + 2]); };
        // ---- split_t_junctions: per edge, its hits in vertex order, put in
        // the reference's cell-scan order, std::sort by t, chained.  The hit
        // vertices of an edge are distinct and never its ends, so edge e
        // becomes exactly hits(e) + 1 edges: the output offsets are a scan of
        // the row popcounts, and the lanes split their own edges (each lane
        // sorts its edge's hits in its own slice of the free adjacency arrays).
        const int rw = tp_row_words(nv);
        int* hv = adjn;                       // E2c  hit vertex (adjn is free until the adjacency below)
        int* hk = adjn + E2c;                 // E2c  scan key
        int ne2 = 0, nh = 0;
        for (int eb0 = 0; eb0 < ne; eb0 += W) {
            const int e = eb0 + ln;
            int nsp = 0;
            if (e < ne)
                for (int w = 0; w < rw; ++w) nsp += __builtin_popcountll(Tq[(int64_t)e * rw + w]);
            int tot_h;
            const int h0 = nh + wp.scan_excl(nsp, tot_h);   // first hit slot of edge e
            const int o0 = ne2 + (h0 - nh) + (e - eb0);        // first output edge of edge e
            const int tot_e = tot_h + (ne - eb0 < W ? ne - eb0 : W);
            if (ne2 + tot_e > E2c) { out.status |= TOPO_E2_OVERFLOW; return out; }
            // the hits of edge e get their parameter and scan key, then are
            // sorted and chained by the edge's own lane when they are at most
            // 16 (the sort is then an insertion sort); longer edges follow
            // below, sorted by the whole wave
            auto split_keys = [&](int a, int b, int hb, int j0, int j1, int dj) {
                const V3 pa = VV(a), pb = VV(b);
                const V3 mn = v3(dmin(pa.x, pb.x) - h, dmin(pa.y, pb.y) - h, dmin(pa.z, pb.z) - h);
                const V3 mx = v3(dmax(pa.x, pb.x) + h, dmax(pa.y, pb.y) + h, dmax(pa.z, pb.z) + h);
                const I3 c0 = cell_of(mn, h), c1 = cell_of(mx, h);
                const int nyz = (c1.y - c0.y + 1) * (c1.z - c0.z + 1);
                for (int j = j0; j < j1; j += dj) {
                    const int v = hv[hb + j];
                    double d2, t;
                    pt_seg(VV(v), pa, pb, d2, t);
                    ht[hb + j] = t;
                    // the reference scans cells ix, iy, iz over [c0, c1]; key = scan index
                    hk[hb + j] = (vcq[3 * v] - c0.x) * nyz + (vcq[3 * v + 1] - c0.y) * (c1.z - c0.z + 1) +
                                 (vcq[3 * v + 2] - c0.z);
                }
            };
            auto split_sort = [&](int hb, int n) {
                SplitArr sa; sa.t = SArr<double>{ht + hb, 1}; sa.v = SArr<int>{hv + hb, 1}; sa.k = SArr<int>{hk + hb, 1};
                SplitByScan byscan;
                for (int i2 = 1; i2 < n; ++i2) {
                    Split x = sa.get(i2);
                    int j2 = i2 - 1;
                    while (j2 >= 0 && byscan(x, sa.get(j2))) { sa.set(j2 + 1, sa.get(j2)); --j2; }
                    sa.set(j2 + 1, x);
                }
                SplitByT byt;
                ref_sort(sa, n, byt);
            };
            constexpr int LANE_SORT = 16;
            if (e < ne) {
                const int a = Eq[2 * e], b = Eq[2 * e + 1];
                int k = 0;
                for (int w = 0; w < rw; ++w)
                    for (uint64_t bits = Tq[(int64_t)e * rw + w]; bits; bits &= bits - 1) hv[h0 + k++] = (w << 6) + ctz64(bits);
                if (nsp <= LANE_SORT) {
                    if (nsp > 0) {
                        split_keys(a, b, h0, 0, nsp, 1);
                        split_sort(h0, nsp);
                    }
                    int prevv = a;
                    for (int j = 0; j < nsp; ++j) {
                        const int v = hv[h0 + j];
                        E2[2 * (o0 + j)] = prevv; E2[2 * (o0 + j) + 1] = v;
                        prevv = v;
                    }
                    E2[2 * (o0 + nsp)] = prevv; E2[2 * (o0 + nsp) + 1] = b;
                }
            }
            wp.sync();
            for (uint64_t big = wp.ballot(e < ne && nsp > LANE_SORT); big; big &= big - 1) {
                const int l = ctz64(big);
                const int ee = wp.readlane(e, l), hb = wp.readlane(h0, l), ob = wp.readlane(o0, l);
                const int n = wp.readlane(nsp, l);
                const int a = Eq[2 * ee], b = Eq[2 * ee + 1];
                split_keys(a, b, hb, ln, n, W);
                wp.sync();
                split_sort(hb, n);
                wp.sync();
                for (int j = ln; j <= n; j += W) {
                    E2[2 * (ob + j)] = j == 0 ? a : hv[hb + j - 1];
                    E2[2 * (ob + j) + 1] = j < n ? hv[hb + j] : b;
                }
                wp.sync();
            }
            ne2 += tot_e;
            nh += tot_h;
        }
        wp.sync();
        wp.mark(0);
        // ---- extract_loops: degrees, adjacency offsets and the adjacency of
        // each vertex in edge order (the reference's append order), lane-parallel
        wp.sync();
        for (int v = ln; v < nv; v += W) deg[v] = 0;
        wp.sync();
        for (int e = ln; e < ne2; e += W) {
            wp.atomic_add(&deg[E2[2 * e]], 1);
            wp.atomic_add(&deg[E2[2 * e + 1]], 1);
        }
        wp.sync();
        {
            int acc = 0;
            for (int vb = 0; vb < nv; vb += W) {
                const int v = vb + ln;
                const int c = v < nv ? deg[v] : 0;
                int tot;
                const int off = wp.scan_excl(c, tot);
                if (v < nv) { aoff[v] = acc + off; afill[v] = acc + off; }
                acc += tot;
            }
        }
        wp.sync();
        // slots 2e (end a) and 2e + 1 (end b) in order: per chunk of W slots,
        // each slot's rank among the chunk's earlier slots of the same vertex
        // (ballots over the bits of the vertex id), then the lowest such lane
        // advances the vertex's cursor
        for (int kb = 0; kb < 2 * ne2; kb += W) {
            const int k = kb + ln;
            const bool act = k < 2 * ne2;
            const int e = k >> 1;
            const int v = act ? E2[k] : 0;
            const int o = act ? E2[k ^ 1] : 0;
            uint64_t same = wp.ballot(act);
            for (int bit = 0; (1 << bit) < nv; ++bit) {
                const bool on = ((v >> bit) & 1) != 0;
                const uint64_t bm = wp.ballot(act && on);
                same &= on ? bm : ~bm;
            }
            const uint64_t below = same & ((1ull << ln) - 1ull);
            if (act) {
                const int pos = afill[v] + (int)__builtin_popcountll(below);
                adjn[pos] = o;
                adje[pos] = e;
            }
            wp.sync();
            if (act && below == 0ull) afill[v] += (int)__builtin_popcountll(same);
            wp.sync();
        }
        // degree-2 vertices (most of a loop's): their two (edge, neighbour)
        // entries in adjacency order, in E2's place (E2 is spent); -1: another degree
        int* pk = E2;
        for (int v = ln; v < nv; v += W) {
            if (deg[v] == 2) {
                const int o = aoff[v];
                pk[4 * v] = adje[o]; pk[4 * v + 1] = adjn[o]; pk[4 * v + 2] = adje[o + 1]; pk[4 * v + 3] = adjn[o + 1];
            } else {
                pk[4 * v] = -1;
            }
        }
        // the used flags of the edges, also as bits in the lanes' registers
        // (edge e: bit e & 63 of lane e >> 6) when they fit, for the walks
        const bool regbits = W == 64 && ne2 <= 64 * W;
        uint64_t usedw = 0;
        auto set_used = [&](int e) {
            used[e] = 1;
            if (regbits && ln == (e >> 6)) usedw |= 1ull << (e & 63);
        };
        // start order: the sort runs on (key, vertex) with the keys in afill
        // (its cursors are spent); same permutation as sorting by StartOrder
        for (int e = ln; e < ne2; e += W) used[e] = 0;
        for (int v = ln; v < nv; v += W) {
            sord[v] = v;
            afill[v] = start_key(deg[v]);
        }
        wp.sync();
        wp.mark(1);
        KeyIdArr ka; ka.k = afill; ka.v = sord;
        ref_sort(ka, nv, ByKey{});
        wp.mark(2);
        int nloops_here = 0;
        // start vertices in order, skipping those whose edges are all used;
        // the next one with an unused edge is found W vertices at a time
        for (int si = 0; si < nv; ++si) {
            {
                int found = nv;
                for (int b0 = si; b0 < nv; b0 += W) {
                    const int s2 = b0 + ln;
                    bool has = false;
                    if (s2 < nv) {
                        const int v2 = sord[s2];
                        for (int m = aoff[v2]; m < aoff[v2] + deg[v2] && !has; ++m) has = used[adje[m]] == 0;
                    }
                    const uint64_t bm = wp.ballot(has);
                    if (bm) { found = b0 + ctz64(bm); break; }
                }
                wp.mark(4);
                if (found >= nv) break;
                si = found;
            }
            const int sv = sord[si];
            for (int k = aoff[sv]; k < aoff[sv] + deg[sv]; ++k) {
                const int ei = adje[k];
                if (used[ei]) continue;
                int plen = 0;
                path[plen++] = sv;
                int prev = sv, cur = adjn[k], ein = ei;   // ein: the edge cur was reached by
                set_used(ei);
                path[plen++] = cur;
                int steps = 0;
                bool closed = false;
                while (steps < 100000) {
                    steps += 1;
                    if (cur == sv) { closed = plen > 1; break; }
                    // the next edge: the first unused one of cur (adjacency
                    // order) that does not lead back to prev, else the first
                    // unused one
                    int pe = 0, pn = 0;
                    const int x0 = pk[4 * cur];
                    if (x0 >= 0) {
                        // degree 2: ein is used, so the other edge, if unused
                        const int n0 = pk[4 * cur + 1], x1 = pk[4 * cur + 2], n1 = pk[4 * cur + 3];
                        pe = x0 == ein ? x1 : x0;
                        pn = x0 == ein ? n1 : n0;
                        const bool u = regbits ? ((wp.readlane64(usedw, pe >> 6) >> (pe & 63)) & 1ull) != 0ull
                                               : used[pe] != 0;
                        if (u) break;
                    } else {
                        // the lanes take W entries at once
                        const int co = aoff[cur], cd = deg[cur];
                        int pick = -1, first = -1, fe = 0, fn = 0;
                        for (int c0 = 0; c0 < cd && pick < 0; c0 += W) {
                            const bool in = c0 + ln < cd;
                            const int m = co + c0 + ln;
                            const int e = in ? adje[m] : 0, nb = in ? adjn[m] : 0;
                            const bool un = in && used[e] == 0;
                            const uint64_t um = wp.ballot(un), pm = wp.ballot(un && nb != prev);
                            if (first < 0 && um) {
                                const int l = ctz64(um);
                                first = c0 + l;
                                fe = wp.readlane(e, l); fn = wp.readlane(nb, l);
                            }
                            if (pm) {
                                const int l = ctz64(pm);
                                pick = c0 + l;
                                pe = wp.readlane(e, l); pn = wp.readlane(nb, l);
                            }
                        }
                        if (first < 0) break;
                        if (pick < 0) { pe = fe; pn = fn; }
                    }
                    set_used(pe);
                    prev = cur;
                    cur = pn;
                    ein = pe;
                    if (plen > E2c) { out.status |= TOPO_E2_OVERFLOW; return out; }
                    path[plen++] = cur;
                }
                wp.mark(5);
                wp.count(plen);
                if (!closed) continue;
                --plen;  // drop the repeated start vertex
                const int raw = out.loops_raw + nloops_here;
                ++nloops_here;
                if (plen < 3) continue;
                // accumulate_Sn_Gn_from_polyline
                V3 Sn = v3(0.0, 0.0, 0.0), Gn = v3(0.0, 0.0, 0.0);
                auto term = [&](const V3& p0, const V3& p1) {
                    Sn = add(Sn, mul(cross(p0, p1), 0.5));
                    const V3 dx = sub(p1, p0);
                    const double s = dot(p0, p1) + (1.0 / 3.0) * dot(dx, dx);
                    Gn = add(Gn, mul(dx, -(1.0 / 3.0) * s));
                };
                int m = 0;
                for (; m + 4 <= plen; m += 4) {   // five vertices' loads in flight, summed in path order
                    const int i0 = path[m], i1 = path[m + 1], i2 = path[m + 2], i3 = path[m + 3];
                    const int i4 = path[m + 4 < plen ? m + 4 : 0];
                    const V3 q0 = VV(i0), q1 = VV(i1), q2 = VV(i2), q3 = VV(i3), q4 = VV(i4);
                    term(q0, q1); term(q1, q2); term(q2, q3); term(q3, q4);
                }
                for (; m < plen; ++m) term(VV(path[m]), VV(path[(m + 1) % plen]));
                // contact_point_xc0
                const double area = norm(Sn);
                if (area < 1e-14) continue;
                V3 nA = mul(Sn, -1.0 / area);
                const V3 xc = dvd(cross(nA, Gn), area);
                if (area < area_eps) continue;
                if (dot(nA, dirAB) < 0.0) nA = mul(nA, -1.0);
                wp.mark(6);
                if (VJ) {
                    tp_loop_jacobian(wp, path, plen, Vs, VJ + 36 * (int64_t)cr[TPC_VB], jac);
                    wp.sync();
                }
                out.status |= sink(out.loops, q, raw, Sn, Gn, nA, xc, area, plen, SArr<int>{path, 1},
                                   SArr<double>{Vs, 1}, VJ ? jac : nullptr);
                out.loops += 1;
                wp.mark(7);
            }
        }
        out.loops_raw += nloops_here;
        wp.sync();
        wp.mark(3);
    }
    return out;
}

// Scratch of the device walks, including their copies of the pair's inputs
// (walk B: the segments, 6S doubles; walk D: edges 2S and vertex cells 6S ints).
ZDX_HD inline int64_t tp_walk_b_ints(int S, int R) { return tp_b_ints(S) + (int64_t)S * (R + 1); }   // + the rows
ZDX_HD inline int64_t tp_walk_b_dbls(int) { return 0; }
ZDX_HD inline int64_t tp_walk_d_ints(int S) { return tp_d_ints(S) + 8 * (int64_t)S; }
ZDX_HD inline int64_t tp_walk_d_dbls(int S) { return tp_d_dbls(S) + 72; }   // + one loop's sensitivities
ZDX_HD inline int64_t tp_walk_ints(int S, int R) {
    const int64_t b = tp_walk_b_ints(S, R), d = tp_walk_d_ints(S);
    return b > d ? b : d;
}
ZDX_HD inline int64_t tp_walk_dbls(int S) {
    const int64_t b = tp_walk_b_dbls(S), d = tp_walk_d_dbls(S);
    return b > d ? b : d;
}
// bytes of a walk's scratch (ints first, doubles 8-aligned)
ZDX_HD inline int64_t tp_scratch_bytes(int64_t ni, int64_t nd) { return ((ni * 4 + 7) & ~(int64_t)7) + nd * 8; }

// ---------------------------------------------------------------- host
// All phases of one pair in sequence (LaneHost), for host-side tests.
template <class Sink>
inline TopoOut tp_topology_host(const double* seg, int S, double h, double area_eps, bool split, const V3& dirAB,
                                Sink& sink, const double* EJ = nullptr) {
    TopoOut out; out.comps = 0; out.loops_raw = 0; out.loops = 0; out.status = TOPO_OK;
    if (S <= 0) return out;
    LaneHost wp;
    std::vector<int> cell9(9 * (size_t)S);
    for (int s = 0; s < S; ++s) tp_cells(seg, s, h, &cell9[9 * (size_t)s]);
    std::vector<int> adjn((size_t)S), adjl((size_t)S * S);
    for (int i = 0; i < S; ++i) adjn[i] = tp_adj_row(seg, cell9.data(), h, S, i, &adjl[(size_t)i * S], S);
    const int n = 2 * S;
    const int64_t npairs = (int64_t)n * (n - 1) / 2;
    std::vector<uint64_t> snapw((size_t)((npairs + 63) >> 6), 0ull);
    for (int64_t t = 0; t < npairs; ++t) {
        int a, b;
        tp_pair_of(t, a, b);
        if (tp_snap_hit(seg, cell9.data(), h, a, b)) snapw[(size_t)(t >> 6)] |= 1ull << (t & 63);
    }
    std::vector<int> IB((size_t)tp_b_ints(S)), comp(TPC_N * (size_t)S), vcell(3 * (size_t)n), E(2 * (size_t)S);
    std::vector<double> V(3 * (size_t)n);
    int tjn = 0;
    std::vector<double> VJ(EJ ? 36 * (size_t)n : 0), jac(72);
    const int ncomp = tp_components_and_vertices(wp, seg, S, h, split, adjn.data(), adjl.data(), S, snapw.data(),
                                                 IB.data(), comp.data(), V.data(), vcell.data(), E.data(), &tjn, EJ,
                                                 EJ ? VJ.data() : nullptr);
    std::vector<uint64_t> tjw((size_t)tjn + 1, 0ull);
    for (int q = 0; q < ncomp; ++q) {
        const int* cr = &comp[TPC_N * (size_t)q];
        const int rw = tp_row_words(cr[TPC_NV]);
        for (int e = 0; e < cr[TPC_NE]; ++e)
            for (int v = 0; v < cr[TPC_NV]; ++v)
                if (tp_tj_hit(&V[3 * (size_t)cr[TPC_VB]], &vcell[3 * (size_t)cr[TPC_VB]], &E[2 * (size_t)cr[TPC_EB]], h, e,
                              v))
                    tjw[(size_t)cr[TPC_TW] + (size_t)e * rw + (v >> 6)] |= 1ull << (v & 63);
    }
    std::vector<int> ID((size_t)tp_d_ints(S));
    std::vector<double> DD((size_t)tp_d_dbls(S));
    return tp_loops(wp, S, ncomp, h, area_eps, dirAB, comp.data(), V.data(), vcell.data(), E.data(), tjw.data(),
                    ID.data(), DD.data(), sink, EJ ? VJ.data() : nullptr, jac.data());
}

}  // namespace zdx
