/*
 * zdx_oracle.h — CPU FP64 restatement of the ZDEM-X Route-B contact step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * product in zedm-x_amd/: it may be loaded by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and by nothing
 * else.  The product never links or calls it.
 *
 * Every function restates the reference algorithm in src/main.cpp of
 * zhl-handsome/ZEDM-X (file:line cited per function in zdx_oracle.cpp), with
 * the same floating-point operation order, compiled with -ffp-contract=off so
 * that results are bit-identical to the reference built with g++ -O2 on
 * x86-64.  The pinning of each function (golden vectors from the compiled
 * reference, closed forms, invariants) is listed at the top of
 * zdx_oracle.cpp.
 */
#ifndef ZDX_ORACLE_H
#define ZDX_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- mesh preparation (main.cpp:287-497) ---- */
/* Reads an STL (binary or ASCII) into tris (T x 9 doubles). Returns T, or -1
 * on failure; if T > cap only the count is returned. */
int orc_load_stl(const char* path, double* tris, int cap);

/* build_mesh: props = {volume, cx, cy, cz, radius, mean_edge, bbox_diag,
 * I00..I22 (9)} = 16 doubles; out_tris receives the (optionally centred)
 * triangles; returns the unique-vertex count. */
int orc_build_mesh(const double* tris, int T, int center, double* out_tris, double* props);

/* ---- per-pair contact (main.cpp:2061-2219 + polyline.py:655-701) ---- */
typedef struct {
    int64_t tri_checks;
    int32_t segments;
    int32_t comps;
    int32_t loops_raw;       /* every traced loop, before filters (main.cpp:2185) */
    int32_t loops;           /* loops accepted by the filters (main.cpp:2188,2202,2206) */
    int32_t loop_verts;      /* total vertices over accepted loops */
    int32_t overflow;        /* 1 if a caller capacity was exceeded */
} orc_pair_counts;

/* One candidate pair: body-frame triangle soups + poses (quat = w,x,y,z).
 * mean_edge_a / bbox_diag_a are mesh A's build_mesh values (tol rule
 * main.cpp:2118-2121, area_eps 2205).  compat_sort=1 reproduces the
 * reference's std::sort loop start order exactly (it always does here: the
 * oracle calls std::sort itself).  Outputs (each may be NULL):
 *   seg      : segments x 6 (p0, p1) in (ta,tb) order, capacity seg_cap
 *   loop_rec : per accepted loop 28 doubles:
 *              Sn(3) Gn(3) n(3, flipped A->B) xc0(3) area Fn F(3) TA(3) TB(3)
 *              comp nverts raw_loop_index   (F is the force on B; A gets -F)
 *   loop_vtx : vertices of accepted loops, 3 doubles each, capacity vtx_cap
 * force: Fn = 0.5*k*area (minimal-energy model). */
int orc_pair_contact(const double* trisA, int TA, const double* trisB, int TB,
                     const double* posA, const double* quatA,
                     const double* posB, const double* quatB,
                     double mean_edge_a, double bbox_diag_a, int split_contacts,
                     double k_stiff,
                     double* seg, int seg_cap,
                     double* loop_rec, int loop_cap,
                     double* loop_vtx, int vtx_cap,
                     orc_pair_counts* counts);

/* ---- simulation (main.cpp:1872-2505 minus output writers) ---- */
typedef struct orc_sim orc_sim;
/* Parse a zdem_cpu config file (main.cpp:1531-1689) and set up particles,
 * meshes and walls (main.cpp:1887-2022).  Relative STL paths are resolved
 * against base_dir when it is non-NULL.  NULL on failure. */
orc_sim* orc_sim_from_config(const char* path, const char* base_dir);
void orc_sim_destroy(orc_sim* s);
int orc_sim_num_particles(const orc_sim* s);
/* Per particle 16 doubles: pos(3) quat(4, w x y z) vel(3) omega(3) L(3). */
void orc_sim_get_state(const orc_sim* s, double* out16);
void orc_sim_set_state(orc_sim* s, const double* in16);
/* forces/torques/contact counts of the last step */
void orc_sim_get_forces(const orc_sim* s, double* F, double* T, int32_t* counts);
/* Run n steps; counters of the LAST step: {contacts, tri_checks, segments,
 * comps, loops, candidate_pairs, wall_contacts} */
int orc_sim_step(orc_sim* s, int n, int64_t* counters7);
double orc_sim_dt(const orc_sim* s);
/* 1: minimal-energy particle-particle force with stiffness k; 0: the reference force */
void orc_sim_set_force_model(orc_sim* s, int model, double k);

/* ---- primitives exposed for the golden checks ---- */
int orc_tri_tri(const double* triA, const double* triB, double* p0p1);
void orc_sn_gn(const double* pts, int m, double* Sn, double* Gn);
int orc_xc0(const double* Sn, const double* Gn, double* xc0, double* nA, double* area);
/* extract_loops start order (std::sort with the reference comparator) */
void orc_loop_start_order(const int32_t* deg, int n, int32_t* order);
/* SH clip of subject (ns x 2) by clip (nc x 2); returns output count */
int orc_sh_clip(const double* subj, int ns, const double* clip, int nc, double* out, int cap);
/* wall contact of one wall triangle; returns 1 if in contact */
int orc_wall_contact(const double* tris_body, int T, const double* pos, const double* quat,
                     const double* wall_tri, const double* wall_n, double tol,
                     double* area, double* xc, double* n);
/* integrate_particle: state16 = pos quat vel omega L (in/out); Ibinv 9; inv_mass */
void orc_integrate(double* state16, const double* Ibody_inv, double inv_mass,
                   const double* F, const double* T, const double* g, double dt);

#ifdef __cplusplus
}
#endif
#endif
