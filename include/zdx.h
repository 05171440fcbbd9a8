/*
 * zdx.h — C ABI of libzdx.so, the MI355X (gfx950) implementation of the
 * ZDEM-X Route-B polyhedral contact step.
 *
 * The reference (zhl-handsome/ZEDM-X) has no library or FFI boundary: its
 * interfaces are the `zdem_cpu --config` program (src/main.cpp:1691-1708 with
 * the config format of src/main.cpp:1531-1689) and the per-pair Python tool
 * polyline.py (polyline.py:707-834).  Each entry point below names the
 * reference code it replaces.  Conventions:
 *   - plain C types only; every function returns a zdx_status;
 *   - the caller owns every host buffer; a context owns device memory and its
 *     HIP streams; calls on one context must not overlap (distinct contexts may
 *     be used from different threads); calls return when results are on the host;
 *   - quaternions are (w, x, y, z); vectors are 3 doubles; all values FP64;
 *   - a ZDX_CAPACITY return fills the required sizes so the caller can retry;
 *   - device buffers grow on demand (a step that overflows one is run again
 *     from the same state); with the environment variable
 *     ZDX_START_CAPACITY=tiny, contexts and simulations created meanwhile
 *     start every such buffer at a few entries (tests of those paths).
 */
#ifndef ZDX_H
#define ZDX_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t zdx_status;
enum {
    ZDX_OK = 0,
    ZDX_BAD_ARG = 1,
    ZDX_IO = 2,
    ZDX_CAPACITY = 3,
    ZDX_HIP_ERROR = 4,
    ZDX_INTERNAL = 5
};

typedef struct zdx_ctx zdx_ctx;
typedef struct zdx_sim zdx_sim;

/* ---------------------------------------------------------------- context */
/* Opens device `device` (HIP ordinal) and creates the context's streams. */
zdx_status zdx_create(zdx_ctx** out, int32_t device);
/* Destroy every simulation created on the context first. */
void zdx_destroy(zdx_ctx* ctx);
/* Message of the last failing call on this context (static storage per ctx). */
const char* zdx_last_error(const zdx_ctx* ctx);

/* ---------------------------------------------------------------- meshes */
/* load_stl, src/main.cpp:308-371 (binary when 84 + 50 n == file size, else
 * ASCII).  Writes up to cap triangles (9 doubles each) and the count to *T. */
zdx_status zdx_load_stl(const char* path, double* tris, int32_t cap, int32_t* T);

/* build_mesh, src/main.cpp:445-497: registers a triangle soup (T x 9 doubles,
 * already scaled) with the context, centring it at its centroid when center
 * != 0, and uploads it.  props (may be NULL) receives 16 doubles:
 * volume, centroid(3), radius, mean_edge, bbox_diag, inertia_unit (9, row-major). */
zdx_status zdx_mesh_register(zdx_ctx* ctx, const double* tris, int32_t T, int32_t center, int32_t* mesh_id,
                             double* props);

/* ------------------------------------------------------ per-pair contact */
/* One contact evaluation of two registered meshes at given poses: the inner
 * body of the reference step, src/main.cpp:2061-2219, with the force of
 * polyline.py:655-701 (Fn = k/2 |Sn|). */
typedef struct {
    int32_t mesh_a, mesh_b;
    double pos_a[3], quat_a[4];
    double pos_b[3], quat_b[4];
} zdx_pair_in;

typedef struct {
    int32_t split_contacts; /* 1: split segments into components (main.cpp:2154-2162) */
    int32_t mode;           /* ZDX_MODE_PLAIN or ZDX_MODE_BATCHED */
    double k;               /* stiffness of Fn = 0.5 k |Sn| */
    int32_t sensitivities;  /* 1: fill zdx_pair_out.loop_jac (no reference counterpart) */
    int32_t reserved;
    double tol;             /* > 0: snap / adjacency tolerance of every pair (polyline.py --tol);
                               0: mesh A's mean edge * 0.1 (main.cpp:2118-2121) */
} zdx_contact_opts;

enum { ZDX_MODE_PLAIN = 0, ZDX_MODE_BATCHED = 1 };

/* Record of one accepted contact loop (28 doubles):
 *   [0:3] Sn  [3:6] Gn  [6:9] n (unit, oriented A->B)  [9:12] xc0  [12] area=|Sn|
 *   [13] Fn  [14:17] F (force on B; A receives -F)  [17:20] torque on A
 *   [20:23] torque on B  [23] component index  [24] vertex count
 *   [25] raw loop index within the pair  [26:28] reserved (0) */
#define ZDX_LOOP_REC 28

typedef struct {
    /* per pair (n_pairs entries each; caller-owned, may be NULL except loop_offset) */
    int64_t* tri_checks;   /* main.cpp:2129 counter */
    int32_t* n_segments;
    int32_t* n_comps;
    int32_t* n_loops_raw;
    int32_t* loop_offset;  /* n_pairs + 1: accepted loops of pair p are [loop_offset[p], loop_offset[p+1]) */
    int32_t* seg_offset;   /* n_pairs + 1, only when segments != NULL */
    /* per accepted loop */
    double* loop_rec;      /* loop_cap x ZDX_LOOP_REC */
    int64_t loop_cap;
    int32_t* vtx_offset;   /* loop_cap + 1, only when loop_vtx != NULL */
    double* loop_vtx;      /* vtx_cap x 3: loop vertices in walk order */
    int64_t vtx_cap;
    double* segments;      /* seg_cap x 6 (p0, p1) in (ta, tb) order */
    int64_t seg_cap;
    /* out: totals (also filled on ZDX_CAPACITY) */
    int64_t n_loops_total, n_vtx_total, n_seg_total;
    /* per accepted loop, with opts->sensitivities: loop_cap x 72 doubles,
     * dSn/dxi (3 x 12, row-major) then dGn/dxi (3 x 12), where xi = (dA, wA,
     * dB, wB) is a rigid motion of the pair's bodies: body X moves its points
     * x by dX + wX x (x - posX).  The loop's topology is held fixed. */
    double* loop_jac;
} zdx_pair_out;

zdx_status zdx_contact_pairs(zdx_ctx* ctx, const zdx_pair_in* pairs, int32_t n_pairs, const zdx_contact_opts* opts,
                             zdx_pair_out* out);

/* zdx_contact_pairs without any device allocation (no counterpart in the
 * reference): for callers that cannot allocate -- code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc.  Same
 * arguments, results and errors as zdx_contact_pairs (the same kernels run),
 * but every device buffer of the call comes from one 256 MiB array per GPU
 * that the library holds for the life of the process; a call that does not
 * fit in it returns ZDX_CAPACITY (use zdx_contact_pairs), and meshes
 * registered since the last allocating call must first be uploaded by one
 * zdx_contact_pairs call (else ZDX_BAD_ARG).  The caller's host buffers are
 * read and written as by zdx_contact_pairs.  All contexts on one GPU share
 * that array: calls that overlap in time on one GPU race and must be
 * serialised by the caller. */
zdx_status zdx_contact_pairs_noalloc(zdx_ctx* ctx, const zdx_pair_in* pairs, int32_t n_pairs,
                                     const zdx_contact_opts* opts, zdx_pair_out* out);

/* Device time of the last zdx_contact_pairs call on this context (no
 * reference counterpart), from HIP events on the context's stream: out[0]
 * transform of the bodies, out[1] narrow phase (work ordering, cull, sweep,
 * exact tests, compaction) of the attempt that fitted the buffers, out[2]
 * topology and loop records, all in ms; out[3] narrow-phase attempts (1 when
 * no buffer had to grow).  The host packing and the transfers are not
 * included.  Writes min(n, 4) values. */
zdx_status zdx_contact_times(zdx_ctx* ctx, double* out, int32_t n);

/* Particle-wall contact geometry (compute_wall_contact, src/main.cpp:1205-1258,
 * for each wall triangle the reference's step tests, 2325-2369): for every
 * (body b, wall triangle t) the plane section of the body's mesh clipped by
 * the triangle, its area and contact point, and the wall normal turned
 * towards the body (main.cpp:2363-2367).  The caller applies the reference's
 * bounding-sphere and box pre-tests (2318-2343) when choosing the pairs.
 * wall_tris: n_wall_tris x 9 world coordinates; wall_normals: x 3.  out:
 * n_bodies x n_wall_tris records of ZDX_WALL_REC doubles:
 *   [0] 1 if in contact  [1] area  [2:5] contact point  [5:8] normal
 *   with jac != 0 (no reference counterpart): [8:14] d area / d xi,
 *   [14:32] d point / d xi (3 x 6, row-major), xi = (d, w) a rigid motion of
 *   the body (its points x move by d + w x (x - pos)). */
typedef struct {
    int32_t mesh, reserved;
    double pos[3], quat[4];
} zdx_body_pose;
#define ZDX_WALL_REC 32
zdx_status zdx_wall_contacts(zdx_ctx* ctx, const zdx_body_pose* bodies, int32_t n_bodies, const double* wall_tris,
                             const double* wall_normals, int32_t n_wall_tris, int32_t jac, double* out);

/* ------------------------------------------------------------- simulation */
/* The zdem_cpu time step (src/main.cpp:2034-2468): broad phase, narrow phase,
 * topology, particle-particle and particle-wall forces with tangential
 * history, force/torque accumulation and the rigid-body integrator. */

/* Loads a zdem_cpu config file (same format; unknown keys ignored) and builds
 * the particles and walls as src/main.cpp:1887-2022 does.  Relative STL paths
 * are resolved against base_dir when it is not NULL. */
zdx_status zdx_sim_load_config(zdx_ctx* ctx, const char* path, const char* base_dir, zdx_sim** out);

/* Programmatic set-up with registered meshes (equivalent to a config whose
 * meshes were loaded once per template and scale). */
typedef struct {
    int32_t n_particles;
    const int32_t* mesh_id;     /* per particle (registered with zdx_mesh_register) */
    const double* init;         /* per particle 13 doubles: pos(3) quat(4) vel(3) omega(3) */
    const double* material;     /* per particle 6 doubles: density young poisson mu restitution (unused) */
    int32_t n_walls;
    const int32_t* wall_mesh_id;
    const double* wall_pose;    /* per wall 7 doubles: pos(3) quat(4) */
    const double* wall_material;/* per wall 2 doubles: mu restitution */
    double gravity[3];
    double dt;
    int32_t split_contacts;
} zdx_sim_desc;
zdx_status zdx_sim_create(zdx_ctx* ctx, const zdx_sim_desc* desc, zdx_sim** out);
void zdx_sim_destroy(zdx_sim* sim);
int32_t zdx_sim_num_particles(const zdx_sim* sim);

/* Step counters (int64 each), as printed on the reference's step line
 * (main.cpp:2489-2499) plus broad-phase and wall totals. */
enum {
    ZDX_CTR_CONTACTS = 0,
    ZDX_CTR_TRI_CHECKS = 1,
    ZDX_CTR_SEGMENTS = 2,
    ZDX_CTR_COMPS = 3,
    ZDX_CTR_LOOPS = 4,
    ZDX_CTR_PAIRS = 5,        /* broad-phase candidate pairs (contact pairs evaluated) */
    ZDX_CTR_WALL_CONTACTS = 6,
    ZDX_CTR_TRI_TESTS = 7,    /* triangle AABB tests actually executed after culling */
    ZDX_CTR_TRI_HITS = 8,     /* triangle pairs that reached the exact intersection test */
    ZDX_CTR_TRI_SCANNED = 9,  /* triangles of both meshes scanned by the per-pair cull (sum of TA + TB) */
    ZDX_CTR_TOPO_PAIRS = 10,  /* pairs with at least one intersection segment */
    ZDX_CTR_SEG_MAX = 11,     /* largest segment count of one pair (zdx_sim_stage_times: max over the steps) */
    ZDX_N_COUNTERS = 12
};
/* Runs n_steps steps in mode ZDX_MODE_PLAIN (one kernel per stage,
 * host-sized launches) or ZDX_MODE_BATCHED (persistent work-queue kernels,
 * whole step captured in a hipGraph).  counters (may be NULL) receives the
 * counters of the last step. */
zdx_status zdx_sim_step(zdx_sim* sim, int32_t n_steps, int32_t mode, int64_t* counters);

/* Per particle 16 doubles: pos(3) quat(4) vel(3) omega(3) L(3). */
zdx_status zdx_sim_get_state(zdx_sim* sim, double* state);
zdx_status zdx_sim_set_state(zdx_sim* sim, const double* state);
/* Force, torque and contact count of each particle from the last step
 * (main.cpp:2027-2029); any pointer may be NULL. */
zdx_status zdx_sim_get_forces(zdx_sim* sim, double* F, double* T, int32_t* contact_counts);
/* Candidate pairs of the last completed step: the broad phase's pairs (i < j,
 * ascending by i then j) of the state that step started from, under the
 * reference's bounding-sphere test (src/main.cpp:2049-2059, rejected when
 * dist2 > (ri + rj)^2).  *n receives the count (0 before the first step);
 * ZDX_CAPACITY when it exceeds cap. */
zdx_status zdx_sim_get_pairs(zdx_sim* sim, int32_t* pi, int32_t* pj, int64_t cap, int64_t* n);
/* write_vtk_particles, src/main.cpp:1710-1840: legacy ASCII VTK polydata of
 * every particle's world triangles with id, mass, radius, contact_count,
 * velocity, omega, force, torque and orientation, byte-compatible with the
 * reference's file for equal state. */
zdx_status zdx_sim_write_vtk(zdx_sim* sim, const char* path);
/* write_particle_state_txt, src/main.cpp:1842-1870: particle blocks (stl,
 * pos, vel, quat, omega, scale, material) that a config file can include.
 * stl and scale are the config's; sims made with zdx_sim_create write an
 * empty stl and scale 1. */
zdx_status zdx_sim_write_state_txt(zdx_sim* sim, const char* path);
/* Exact binary checkpoint of state and tangential histories (the
 * reference's only restart file, write_particle_state_txt main.cpp:1842-1870,
 * keeps 9 decimals and no history). */
zdx_status zdx_sim_checkpoint(zdx_sim* sim, const char* path);
zdx_status zdx_sim_restore(zdx_sim* sim, const char* path);
/* Stage timing and counters accumulated over the steps since the previous
 * call, then reset.  Each stage's kernels stamp their span (earliest
 * workgroup start, latest wave end) on the device's constant-rate clock, in
 * both modes and inside the captured graph.  out[0..5]:
 * mean ms per step of transform, broad phase, narrow phase, topology+forces,
 * walls, fold+integrate+counters; out[6]: steps; out[7..7+ZDX_N_COUNTERS):
 * counter sums (ZDX_CTR_SEG_MAX: the maximum); out[7+ZDX_N_COUNTERS]: mean ms
 * of the narrow-phase kernels alone (without the work ordering);
 * out[8+ZDX_N_COUNTERS]: steps run again after a buffer growth;
 * out[9+ZDX_N_COUNTERS]: captures of the batched step's graphs (a growth or a
 * parameter change invalidates them).  The walls run on a second stream
 * alongside the narrow phase and the topology, so the stages can add up to
 * more than the step.  Writes min(n, 10 + ZDX_N_COUNTERS) values. */
zdx_status zdx_sim_stage_times(zdx_sim* sim, double* out, int32_t n);
/* Particle-particle force model of the time step.  ZDX_FORCE_ZDEM (default):
 * the reference's force, src/main.cpp:2220-2287.  ZDX_FORCE_MINIMAL: the
 * minimal-energy force of polyline.py:655-701, Fn = k/2 |Sn| along the A->B
 * normal, applied at the loop centroid, with no damping and no tangential
 * force.  ZDX_FORCE_ZDEM_DAMPED: the reference's force plus the normal
 * damping it computes and leaves commented out (main.cpp:2244-2248,
 * 2258-2260): fn += -cn vn while the bodies approach.  Walls keep the
 * reference's wall force in every model.  In a config file: "force_model
 * minimal" with "minimal_k <k>", or "force_model zdem_damped". */
enum { ZDX_FORCE_ZDEM = 0, ZDX_FORCE_MINIMAL = 1, ZDX_FORCE_ZDEM_DAMPED = 2 };
zdx_status zdx_sim_set_force_model(zdx_sim* sim, int32_t model, double k);
/* Narrow-phase launch shape (no reference counterpart).  The narrow phase
 * runs every pair on one wave each and hands the pairs too large for one wave
 * to a workgroup kernel; np_shape encodes that kernel's workgroup size and
 * triangle-pair tests per thread of one sweep round as threads * 100 +
 * tests: one of 6404, 6408, 12804, 12808, 25601, 25602, 25604, 25608, 51202,
 * 51204 (the compiled kernels); any other value returns ZDX_BAD_ARG.
 * np_groups_per_cu (1..16) sets its persistent grid to that many workgroups
 * per CU.  0 keeps the current value of either.  The autotuner
 * (zdx_sim_autotune) times every compiled shape, and the most float-box
 * tests a pair may have on the one-wave kernel (4096 or 16384), on the
 * current step's pair set and keeps the fastest; it reports the shape in the
 * same encoding. */
zdx_status zdx_sim_set_tiles(zdx_sim* sim, int32_t np_shape, int32_t np_groups_per_cu);
zdx_status zdx_sim_autotune(zdx_sim* sim, int32_t* np_shape, int32_t* np_groups_per_cu);

#ifdef __cplusplus
}
#endif
#endif /* ZDX_H */
