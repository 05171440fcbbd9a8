// zdx_internal.h — data layouts shared by the host code and the HIP kernels.
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

namespace zdx {

// ---------------------------------------------------------------- host mesh
// Result of build_mesh (src/main.cpp:445-497) plus the indexed form the GPU
// uses: corners of the (centred) triangles that are bitwise equal share one
// vertex slot, so each particle's vertices are rotated once per step and the
// triangle corners read back the identical doubles.
struct HostMesh {
    std::vector<double> tris;      // T x 9, centred when requested
    std::vector<double> verts;     // unique corners (bitwise), 3 doubles each
    std::vector<int32_t> tri_idx;  // T x 3 indices into verts
    double volume = 0, radius = 0, mean_edge = 0, bbox_diag = 0;
    double center[3] = {0, 0, 0};
    double inertia[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    int n_unique_ref = 0;          // the reference's unique_vertices count (tolerance merge)
    std::vector<int32_t> suspect;  // triangles whose world-frame plane test is not certain
};

bool load_stl(const std::string& path, std::vector<double>& tris);
HostMesh build_mesh(const std::vector<double>& tris, bool center);

// config (src/main.cpp:1512-1689)
struct ParticleInit {
    std::string stl;
    double pos[3] = {0, 0, 0}, vel[3] = {0, 0, 0}, omega[3] = {0, 0, 0};
    double quat[4] = {1, 0, 0, 0};
    double scale = 0.0, density = 1.0, young = 1e7, poisson = 0.25, mu = 0.5, rest = 0.5;
};
struct WallInit {
    std::string stl;
    double pos[3] = {0, 0, 0};
    double quat[4] = {1, 0, 0, 0};
    double scale = 1.0, mu = 0.5, rest = 0.5;
};
struct SimConfig {
    int n = 2, steps = 1;
    double dt = 1e-4, spacing = 2.5;
    double gravity[3] = {0.0, -9.81, 0.0};
    bool split = true, center = true, contact_debug = false;
    std::string stl, vtk_prefix = "particles", output_dir = "output";
    int output_interval = 1;
    double v0[3] = {0, 0, 0}, v1[3] = {0, 0, 0};
    int force_model = 0;        // build extension: "force_model minimal" / "zdem_damped" (the reference
                                // ignores the key)
    double k_minimal = 0.0;     // "minimal_k <k>"
    std::vector<ParticleInit> parts;
    std::vector<WallInit> walls;
};
bool parse_config(const std::string& path, SimConfig& cfg);

// The particles and walls a config describes (src/main.cpp:1887-2022): one
// mesh per (stl path, scale) as the reference's `path|to_string(scale)` cache,
// the per-particle inits (the n/stl/spacing/v0/v1 form expanded as the
// reference does), and the walls whose STL loads.  Relative paths are
// resolved against base_dir when it is not empty.
struct ConfigScene {
    std::vector<HostMesh> meshes;
    std::vector<int32_t> mesh_of;          // per particle, index into meshes
    std::vector<ParticleInit> inits;       // per particle, as configured
    std::vector<HostMesh> wall_meshes;     // per loaded wall (scaled, centred like particles)
    std::vector<WallInit> walls;           // per loaded wall
};
bool build_scene(const SimConfig& cfg, const std::string& base_dir, ConfigScene& scene, std::string& err);

// per-step output files (src/main.cpp:1710-1870); mesh[i] is particle i's mesh,
// state rows are 16 doubles (pos quat vel omega L)
bool write_vtk_particles(const std::string& path, int n, const HostMesh* const* mesh, const double* state,
                         const double* mass, const double* radius, const double* F, const double* T,
                         const int32_t* counts);
bool write_particle_state_txt(const std::string& path, int n, const double* state, const ParticleInit* init);

// polyline.py's outputs (polyline.py:544-649): Python's repr of a float (the
// text of f"{x}": shortest round-trip digits, fixed notation for decimal
// exponents -4 <= e < 16 with ".0" on integral values, else d.ddde+XX), the
// transformed-mesh POLYDATA (points merged within 1e-9 in order of first
// use) and the contact-loop POLYDATA (one closed LINES cell per loop, one
// VERTICES cell per contact point, CELL_DATA contact_id).
std::string py_float_repr(double x);
bool write_vtk_polydata_tris(const std::string& path, const double* tris, int64_t T, const std::string& comment);
struct V3;
// contact_points[k] with has_point[k] == 0 stands for polyline.py's None (counted, not written)
bool write_vtk_polydata_polylines(const std::string& path, const std::vector<std::vector<V3>>& loops,
                                  const std::vector<V3>& contact_points, const std::vector<char>& has_point,
                                  const std::vector<int>& loop_contact_ids);

// polyline.py's transform_points (polyline.py:36-52) of n points (3 doubles
// each): q divided by its norm (identity below 1e-30), v + w t + qv x t with
// t = 2 qv x v, plus the translation.
void polyline_transform_points(const double* in, int64_t n, const double* t, const double* q, double* out);

// polyline.py's contact of one closed loop of m points (polyline.py:193-218,
// 676-679): Sn, Gn summed edge by edge from the first point, area = |Sn|,
// nA = -Sn / area, xc0 = (nA x Gn) / area (none when area < 1e-14, then nA = 0),
// Fn = 0.5 k area, F_A = Fn nA.  Dots and norms are numpy's (a running fused
// multiply-add sum), products and differences are rounded one by one.
struct PolylineContact {
    double Sn[3], Gn[3], area, nA[3], xc0[3], Fn, FA[3];
    bool has_xc0;
};
PolylineContact polyline_contact(const double* pts, int64_t m, double k);
// polyline.py's loop filter (polyline.py:798-806): kept when it has at least
// 3 points and the diagonal of its bounding box (numpy's norm) is >= 2 tol.
bool polyline_keeps_loop(const double* pts, int64_t m, double tol);

// per-particle constants derived at set-up (src/main.cpp:1955-1985)
struct ParticleConst {
    double mass, inv_mass, Ibody[9], Ibinv[9], radius, eqr, young, poisson, mu, rest;
    int32_t mesh;
};
ParticleConst particle_const(const HostMesh& m, int mesh_id, double density, double young, double poisson,
                             double mu, double rest);
// L = I_world * omega (src/main.cpp:1985)
void initial_L(const ParticleConst& c, const double* quat, const double* omega, double* L);

// ---------------------------------------------------------------- device views
// Registered meshes, concatenated.
struct DMeshes {
    const double* verts;     // body-frame vertices
    const int32_t* tri_idx;  // local vertex indices, 3 per triangle
    const int32_t* vert_off; // per mesh
    const int32_t* nverts;
    const int32_t* tri_off;
    const int32_t* ntris;
    const double* tol;       // per mesh: mean_edge*0.1 rule (main.cpp:2118-2121)
    const double* area_eps;  // per mesh (main.cpp:2205)
    const int32_t* susp_off; // per mesh: first suspect triangle in susp
    const int32_t* nsusp;    // per mesh: triangles whose world plane may degenerate
    const int32_t* susp;     // suspect triangle indices
};

// Bodies whose world geometry is evaluated each step.
struct DBodies {
    int32_t n;
    const int32_t* mesh;     // per body
    const double* state;     // 16 doubles per body (pos quat vel omega L)
    const int64_t* wv_off;   // per body: first world vertex
    const int64_t* wt_off;   // per body: first triangle flag
    double* wv;              // world vertices (3 doubles each)
    int32_t* nondeg;         // per body: non-degenerate triangle count
    double* aabb;            // per body: min(3) max(3) over world vertices
};

struct DMat {                // per particle material / inertia (sim only)
    const double* mass;
    const double* inv_mass;
    const double* Ibinv;     // 9 per particle
    const double* eqr;
    const double* young;
    const double* poisson;
    const double* mu;
    const double* rest;
    const double* radius;
};

// Narrow-phase output arena and per-pair results.
// One pair of the narrow phase in work order, gathered by the work-ordering
// kernel so that the cull starts from a single 128-byte load: pair, bodies,
// triangle / vertex counts and offsets of both meshes, A's non-degenerate
// triangle count and both bodies' outward-rounded float AABBs.
struct NarrowItem {
    int32_t p, i, j, TA, TB, VA, VB, toA, toB, woA, woB, nondegA;
    float boxA[6], boxB[6];
    int32_t pad[8];
};

// The hit arena is cut into shards, each with its own counter: a pair's
// hits go to shard (work item % hit_shards), so that the reservations are
// spread over that many addresses (atomics on one address serialise, at about
// 11.5 ns each on MI355X, and so do atomics on one 128-byte line: each counter
// has a line of its own, HIT_TOP_STRIDE apart).
constexpr int HIT_SHARDS = 64;
constexpr int HIT_TOP_STRIDE = 16;   // in 8-byte counters

struct DPairs {
    int32_t cap;             // capacity of the pair arrays
    const int32_t* n;        // device count of pairs
    const int32_t* pi;       // body A (lower index in the sim)
    const int32_t* pj;       // body B
    const int32_t* order;    // processing order (LPT in batched mode)
    NarrowItem* items;       // with order: the pairs' narrow-phase items in that order
    int64_t* seg_off;        // per pair: first segment in the arena
    int32_t* nseg;           // per pair
    int64_t* tri_checks;     // per pair (reference counter)
    int64_t* tri_tests;      // per pair (executed AABB tests)
    int32_t* tri_hits;       // per pair
    int32_t* tri_scanned;    // per pair: TA + TB
    double* arena;           // segments, 6 doubles each, one slot per hit slot
    int64_t arena_cap;       // in segments (= hit_cap)
    int32_t* topo_list;      // pairs with at least one segment (appended by the narrow phase)
    int32_t* topo_n;         // device count of topo_list
    int32_t* topo_big;       // optional: topo_list indices of pairs with many segments (walked first) ...
    int32_t* topo_small;     // ... and of the others; with their device counts
    int32_t* n_big;
    int32_t* n_small;
    int32_t* status;         // global error bits
    int32_t* prov;           // optional (sensitivity runs): 4 ints per arena segment: ta, tb, endpoint codes
    int32_t* hits;           // hit arena: per triangle pair that passed the float AABB sweep, the six
                             // world-vertex indices of (ta, tb); its slots are also the segment slots
    int32_t* hit_tri;        // optional (sensitivity runs): (ta, tb) per hit
    int64_t hit_cap;         // in hits; shard k holds [k, k + 1) * (hit_cap / hit_shards)
    unsigned long long* hit_top;   // hit_shards counters, HIT_TOP_STRIDE apart: hits each shard's pairs asked for
    int32_t hit_shards;      // 1 .. HIT_SHARDS
    int64_t* hit_off;        // per pair: first hit (count in tri_hits); the pair's segments start there
    int32_t* hit_flag;       // per hit: 1 when the exact test produced a segment
    int32_t* heavy;          // narrow phase: work items left by the one-wave kernel for the workgroup kernel
};

enum : int32_t {
    ST_ARENA = 1,        // (no longer raised: segment slots are the hit slots)
    ST_STAGE = 2,        // a pair produced more segments than the staging capacity
    ST_TOPO = 4,         // topology scratch or loop capacity exceeded
    ST_PAIRS = 8,        // broad-phase pair capacity exceeded
    ST_NEIGH = 16,       // (no longer raised: the broad phase keeps every partner of a particle)
    ST_WALL = 32,        // wall section capacity exceeded
    ST_GRID = 64,        // broad-phase grid could not be built (non-finite positions)
    ST_HITS = 128        // hit arena full
};

}  // namespace zdx
