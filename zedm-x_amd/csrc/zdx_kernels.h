// zdx_kernels.h — host-side launchers of the HIP kernels (zdx_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "zdx_internal.h"

namespace zdx {

// Broad-phase grid: origin, cell size and dimensions chosen on the device.
struct DGrid {
    double ox, oy, oz, h;
    int32_t nx, ny, nz;
    int32_t ncells;
};

constexpr int BP_PARTS = 256;   // workgroups of the bounds reduction

struct BroadPhaseBufs {
    int32_t n;                 // particles
    const double* state;       // 16 per particle
    const double* radius;
    double* part;              // BP_PARTS x 7 partial bounds (min xyz, max xyz, max radius)
    DGrid* grid;               // device
    int32_t max_cells;
    int32_t* cell_of;          // per particle
    int32_t* cell_count;       // max_cells
    int32_t* cell_start;       // max_cells + 1
    int32_t* sorted;           // per particle
    double* spos;              // per cell-order slot: x, y, z, radius
    int32_t* cnt_up;           // per particle
    int32_t* cnt_all;          // per particle
    int32_t* nbr;              // per cell-order slot: the first 16 partners found by the count pass
    int32_t* pair_off;         // n + 1
    int32_t* adj_off;          // n + 1
    int32_t* pi;               // pair capacity
    int32_t* pj;
    int64_t* pkey;
    int32_t pair_cap;
    int32_t* adj_j;            // adjacency capacity
    int32_t* adj_pair;
    int32_t adj_cap;
    int32_t* n_pairs;          // device scalar
    int32_t* n_adj;            // device scalar
    int32_t* scan_tmp;         // scan block sums (>= max_cells/1024 + 2)
    int32_t* status;
};

struct WallBufs {
    int32_t n_walls;
    const double* tris;        // world wall triangles, 9 doubles each
    const double* normals;     // per wall triangle
    const int32_t* tri_off;    // per wall
    const int32_t* ntri;       // per wall
    const double* pos;         // per wall
    const double* radius;      // per wall (mesh radius)
    const double* mu;          // per wall
    int32_t max_wall_tris;     // contributions per (particle, wall)
};

// Per-step history (tangential springs): the previous step's sorted pair
// keys with their final ds and active flags, and this step's output.
struct HistBufs {
    const int64_t* prev_key;
    const double* prev_ds;
    const uint8_t* prev_active;
    const int32_t* prev_n;
    double* cur_ds;            // per pair of this step
    uint8_t* cur_active;
    // walls, dense [particle * n_walls + w]
    const double* wprev_ds;
    const uint8_t* wprev_active;
    double* wcur_ds;
    uint8_t* wcur_active;
};

struct TopoBufs {
    int32_t lmax;              // contact records per pair
    double* contrib;           // pair * lmax * 9  (f, rA, rB)
    int32_t* ncontrib;         // per pair: accepted loops
    int32_t* comps;            // per pair
    int32_t* loops_raw;        // per pair
    int32_t* tstatus;          // per pair
    // phased topology (zdx_topo.h); "item" = index into the pair work list
    int32_t* cell9;            // 9 per arena segment
    int32_t* adj;              // adjacency rows: per pair S counts + S x adj_row entries (bump-allocated)
    int64_t adj_cap;
    int32_t adj_row;           // entries per adjacency row
    uint64_t* snapw;           // snap hit words (bump-allocated)
    int64_t snap_cap;
    uint64_t* tjw;             // T-junction hit words (bump-allocated)
    int64_t tj_cap;
    unsigned long long* top;   // counters: adjacency ints, snap words, T-junction words; longest row; pass-A units
    int32_t* units;            // pass A: extra units (item, unit) of the pairs split into several
    int64_t units_cap;
    int64_t* adj_off;          // per item (-1: did not fit)
    int64_t* snap_off;
    int64_t* tj_off;
    int32_t* ncomp;            // per item (-1: not evaluated)
    int32_t* comp;             // TPC_N per arena segment
    double* V;                 // 2 vertices (6 doubles) per arena segment
    int32_t* vcell;            // 2 vertex cells (6 ints) per arena segment
    int32_t* E;                // 1 edge (2 ints) per arena segment
    int32_t* lab;              // 2 per arena segment: snap cluster (its smallest endpoint) of each endpoint, for
                               // the pairs of the big list (k_tp_snap_cc); first entry -1: not computed
    // sensitivities (API runs with provenance; nullptr otherwise)
    double* ej;                // 2 endpoint Jacobians (72 doubles) per arena segment
    double* vj;                // 2 vertex Jacobians (72 doubles) per arena segment
    double* ljac;              // pair * lmax * 72: dSn/dxi, dGn/dxi per loop
    // walk scratch: LDS of lds_bytes when a pair fits, else per-workgroup slices
    int32_t* scr_i;
    double* scr_d;
    int64_t scr_ni, scr_nd;    // slice sizes (for seg_max segments)
    int64_t nthreads;          // persistent one-wave workgroups of the walks (= slices)
    int64_t pass_groups;       // persistent 256-thread workgroups of the passes
    int32_t seg_max;           // largest segment count the slices hold
    int32_t lds_bytes;
    // API mode outputs
    double* rec;               // pair * lmax * 28
    double* vtx;               // pair * vmax * 3
    int32_t vmax;
    double k_stiff;
    double tol_override;       // > 0: snap / adjacency tolerance for every pair (per-pair API)
    unsigned long long* stamp_alloc;   // clock written when k_tp_alloc starts, or nullptr
};

struct SimParams {
    double g[3];
    double dt;
    int32_t split;
    int32_t force_model;   // 0: the reference's particle-particle force, 1: minimal-energy force,
                           // 2: the reference's force with its dormant normal damping
    double k_minimal;      // stiffness k of the minimal-energy force
};

struct NarrowCfg {
    int32_t block;             // 64, 128, 256 or 512 threads
    int32_t items;             // AABB tests per thread per round
    int32_t groups;            // persistent workgroups (of the workgroup kernel, for the heavy pairs)
    int32_t use_queue;         // 1: atomic work queue (batched)
    int32_t wave_groups;       // workgroups of the one-wave-per-pair kernel (0: 4 per CU's worth of groups)
    int32_t wave_tests;        // most float-box tests of a pair on the one-wave kernel (0: its default)
};

// Writes the device clock into *slot when the stream reaches it (NULL: no-op).
void launch_stamp(unsigned long long* slot, hipStream_t s);
// Small arrays a step clears before its stages run (NULL entries skipped), by
// launch_broadphase's first kernel.
struct StepZero {
    int32_t* i32[8];
    int32_t n32[8];
    unsigned long long* u64[2];
    int32_t n64[2];
};

// What the host reads after every step, in one block of pinned, coherent host
// memory that the last kernel of the step writes: status, counters, stage
// clocks, then seq = 1 (the host clears seq before it launches the step and
// waits for it).
struct StepReport {
    int32_t status;
    int32_t pad;
    unsigned long long ctr[16];
    unsigned long long kt[12];
    unsigned int seq;
    unsigned int pad2;
};
// Copies status, ctr and kt into rep, zeroes ctr for the next step's sums and
// publishes rep->seq.
void launch_report(const int32_t* status, unsigned long long* ctr, const unsigned long long* kt, StepReport* rep,
                   hipStream_t s);
// list/n_list: bodies to transform (device list and count; NULL = all bodies);
// stamp: clock written by the first workgroup at its start
void launch_transform(const DBodies& B, const DMeshes& M, hipStream_t s, const int32_t* list = nullptr,
                      const int32_t* n_list = nullptr, int32_t groups = 0, unsigned long long* stamp = nullptr);
// bodies that are in a candidate pair or near a wall triangle -> list
// (and the particles near a wall triangle -> wlist), in no fixed order;
// zeroed: *n_list and *n_wlist already hold 0
void launch_need_list(int32_t n, const double* state, const double* radius, const int32_t* cnt_all, const WallBufs& W,
                      int32_t* list, int32_t* n_list, int32_t* wlist, int32_t* n_wlist, bool zeroed, hipStream_t s,
                      unsigned long long* stamp = nullptr);
// The same lists built by the broad phase's count pass (its counts and
// n_list, n_wlist zero beforehand): launch_broadphase with need set.
struct NeedLists {
    WallBufs W;
    int32_t* list;
    int32_t* n_list;
    int32_t* wlist;
    int32_t* n_wlist;
};
// kt0 and z: the broad phase's first kernel also does launch_step_begin's work
// (kt0 the step's clock 0, z the arrays to zero), so the step needs no kernel
// of its own for it
void launch_broadphase(const BroadPhaseBufs& bp, hipStream_t s, const NeedLists* need = nullptr,
                       unsigned long long* kt0 = nullptr, const StepZero* z = nullptr);
// Narrow phase: cull + sweep into the hit arena, the exact tests per hit, and
// the per-pair compaction into the segment arena.  seg_cap: most segments one
// pair may store (ST_STAGE beyond); queue: the cull's work counter; resets
// the queue and P.hit_top itself unless the caller did (zeroed).
// queue: NARROW_COUNTERS counters (the wave kernel's work queues, the heavy
// list's front and back counts, the workgroup kernel's queue over it, each on a
// 128-byte line of its own), zeroed here unless `zeroed`.
constexpr int NARROW_COUNTERS = 67 * 32;
void launch_narrow(const DBodies& B, const DMeshes& M, const DPairs& P, const NarrowCfg& cfg, int32_t* list_i,
                   float* list_f, int32_t seg_cap, int32_t tmax, unsigned int* queue, hipStream_t s, bool zeroed = false,
                   unsigned long long* stamp = nullptr);
// LPT work order of the pairs (largest estimated cost first) into order,
// and the narrow-phase items (P.items, when set) in that order; zeroed:
// bucket_cnt and bucket_cur (64 each) already hold 0
// tz: when set, the scatter pass also zeroes the topology's per-pair outputs
// (what launch_topo_zero does, with cur_active)
void launch_lpt_order(const DPairs& P, const DMeshes& M, const DBodies& B, int32_t* order, int32_t* bucket_cnt,
                      int32_t* bucket_cur, int32_t pair_cap, hipStream_t s, bool zeroed = false,
                      unsigned long long* stamp = nullptr, const TopoBufs* tz = nullptr,
                      uint8_t* cur_active = nullptr);
// zero outputs of every pair (the topology kernels then visit only the
// pairs of P.topo_list); cur_active may be NULL
void launch_topo_zero(const DPairs& P, const TopoBufs& T, uint8_t* cur_active, int32_t pair_cap, hipStream_t s,
                      unsigned long long* stamp = nullptr);
// side/ev_fork/ev_join (optional, with P.topo_big): pass A of the many-segment
// pairs runs on `side`, forked from and joined back into s
void launch_topo_sim(const DPairs& P, const DMeshes& M, const DBodies& B, const DMat& mat, const HistBufs& H,
                     const TopoBufs& T, const SimParams& sp, hipStream_t s, bool zeroed = false,
                     hipStream_t side = nullptr, hipEvent_t ev_fork = nullptr, hipEvent_t ev_join = nullptr,
                     hipEvent_t ev_cc = nullptr);
void launch_topo_api(const DPairs& P, const DMeshes& M, const DBodies& B, const TopoBufs& T, int32_t split,
                     hipStream_t s);
// one wave per particle of `list` (particles near a wall triangle); scr holds
// `groups` slices of 3 pts_cap + 4 (pts_cap + 8) doubles
void launch_walls(const DBodies& B, const DMeshes& M, const DMat& mat, const WallBufs& W, const HistBufs& H,
                  double* wcontrib, int32_t* wncontrib, const int32_t* list, const int32_t* n_list, double* scr,
                  int64_t groups, int32_t pts_cap, const SimParams& sp, int32_t* status, hipStream_t s,
                  unsigned long long* stamp = nullptr);
void launch_fold_integrate(int32_t n, const double* state_in, double* state_out, const DMat& mat,
                           const int32_t* adj_off, const int32_t* adj_j, const int32_t* adj_pair,
                           int32_t adj_cap, const TopoBufs& T, int32_t n_walls, const double* wcontrib, const int32_t* wncontrib,
                           int32_t max_wall_tris, double* F, double* Tq, int32_t* counts, const SimParams& sp,
                           const DPairs& P, unsigned long long* counters, hipStream_t s,
                           unsigned long long* stamp_a = nullptr, unsigned long long* stamp_b = nullptr);
// (counters: ZDX_N_COUNTERS int64 summed by the fold's first workgroups from
// the per-pair and per-wall arrays; they must hold zeros, which launch_report
// leaves; nullptr skips them)
void launch_fill_i32(int32_t* p, int32_t v, int64_t n, hipStream_t s);
void launch_fill_u64(unsigned long long* p, unsigned long long v, int64_t n, hipStream_t s);
void launch_iota_pairs(int32_t* pi, int32_t* pj, int32_t n, hipStream_t s);

// Output of the per-pair call packed on the device, so that each array
// reaches the caller in one copy: per-pair loop, vertex and segment offsets
// (exclusive scans), then the accepted loops' records (28 doubles), their
// Jacobians (72, optional), their vertices (optional) and the segments
// (optional) back to back in pair order.  Optional outputs are NULL.
struct ApiPack {
    int32_t n;                 // pairs
    int32_t lmax, vmax;        // per-pair slots of rec / ljac and of vtx
    const int32_t* nloops;     // accepted loops per pair
    const int32_t* nseg;       // segments per pair
    const int64_t* seg_off;    // the pair's first segment in the arena
    const double* rec;         // n * lmax * 28
    const double* ljac;        // n * lmax * 72
    const double* vtx;         // n * vmax * 3
    const double* arena;       // segments, 6 doubles each
    int32_t* loop_off;         // n + 1
    int32_t* seg_out_off;      // n + 1
    int64_t* pvtx_off;         // n + 1: first vertex of each pair's loops
    int64_t* totals;           // loops, vertices, segments
    double* rec_out;           // loop_cap * 28
    double* ljac_out;          // loop_cap * 72
    double* vtx_out;           // vtx_cap * 3
    int32_t* vtx_off_out;      // loop_cap + 1
    double* seg_out;           // seg_cap * 6
    int64_t loop_cap, vtx_cap, seg_cap;   // rows of the *_out arrays: nothing is written past them
};
void launch_api_pack(const ApiPack& k, hipStream_t s);
// Bodies of the per-pair call from the caller's pair records (zdx_pair_in,
// copied to the device as they are, `in`): body 2p is pair p's A, 2p + 1 its
// B.  Writes mesh ids, state rows (pos, quat, the rest 0) and each body's
// first world vertex and triangle (exclusive scans over the meshes' counts,
// 2n + 1 entries: the last holds the totals).
void launch_api_bodies(const void* in, int32_t n, const DMeshes& M, int32_t* mesh, double* state, int64_t* wv_off,
                       int64_t* wt_off, hipStream_t s);
// Wall contacts of bodies against wall triangles (the per-triangle part of
// the reference's particle-wall contact, src/main.cpp:2325-2369, without its
// bounding-sphere and box pre-tests): one wave per (body b, triangle t);
// out[32 (b nt + t) ...]: flag, area, xc(3), normal(3, the wall normal turned
// towards the body's centre, main.cpp:2363-2367), and with jac d area/dxi
// (6) and d xc/dxi (3 x 6) for rigid motions xi = (d, w) of the body.
// scr: per wave 3 cap + 4 (cap + 8) + 28 (cap + 8) doubles and cap + 1 ints
// (in isrc).  status bit 1: a section held more than cap points.
void launch_wall_query(const DBodies& B, const DMeshes& M, int32_t n_bodies, const double* tris, const double* normals,
                       int32_t n_tris, int32_t jac, int32_t cap, double* scr, int32_t* isrc, double* out,
                       int32_t* status, hipStream_t s);
// *out = OR of v[0..n) (out is cleared first)
void launch_or_i32(const int32_t* v, int32_t n, int32_t* out, hipStream_t s);

}  // namespace zdx
