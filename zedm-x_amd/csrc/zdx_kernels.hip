// zdx_kernels.hip — gfx950 kernels of the Route-B contact step.
//
// FP64 throughout; built with -ffp-contract=off so that every expression
// rounds as the reference's CPU build does (see zdx_geom.h).
//
//   k_transform       one wave per body: world vertices (quat_rotate of
//                     each unique body vertex), body AABB, degenerate-triangle
//                     flags and counts
//   broad phase       device-sized uniform grid (cell >= 2 r_max), counting
//                     sort of particles by cell, per-particle neighbour
//                     enumeration, scans, ordered pair list (i < j) and the
//                     per-particle adjacency used by the ordered force fold
//   narrow phase      k_narrow_wave (one wave per pair) and, for the pairs
//                     too large for a wave, k_narrow (persistent workgroups):
//                     cull each mesh's triangles against the other mesh's
//                     box, then a flattened, order-preserving sweep of the
//                     culled product with the float (outward-rounded) AABB
//                     test, hits written in (ta, tb) order; k_narrow_exact:
//                     one thread per hit, the exact FP64 AABB re-test and
//                     tri-tri; k_narrow_compact: one wave per pair, ballot /
//                     prefix-sum compaction in the reference's (ta, tb) order
//   topology          k_tp_alloc (one thread per pair) and k_tp_pass_a
//                     (workgroups over flat S^2 test ranges) for the
//                     quadratic tests; k_tp_snap_cc (one workgroup per pair
//                     of 40 or more segments: its snap clusters); walks B and
//                     D with pass C between (components, snap union-find,
//                     T-junction chains, loops, Sn/Gn/xc0, force) one wave
//                     per pair with scratch in LDS (zdx_topo.h)
//   k_walls           one wave per particle near a wall: plane sections of
//                     every candidate wall triangle, clip, area, force
//   k_fold_integrate  one thread per particle: ordered, atomic-free fold of
//                     the contact records in the reference's summation order,
//                     then the rigid-body integrator; its first workgroups
//                     also sum the step's counters
#include <hip/hip_runtime.h>

#include "../../include/zdx.h"
#include "zdx_geom.h"
#include "zdx_kernels.h"
#include "zdx_topo.h"

namespace zdx {

namespace {

constexpr int WAVE = 64;

#include "zdx_wave.h"

// Stage stamps: one thread writes the device's constant-rate clock into a
// slot; launched on the stream between stages, so consecutive stamps bracket
// a stage (kernels of a stream run in order, also inside a captured graph).
// A per-kernel scheme with an atomic from every wave would serialise tens of
// thousands of same-address atomics per launch.
__global__ void k_stamp(unsigned long long* slot) { *slot = (unsigned long long)wall_clock64(); }

// The end-of-step report into host memory (fine-grained, so the stores reach
// it in order once fenced); seq goes last.
__global__ void __launch_bounds__(64) k_report(const int32_t* status, unsigned long long* ctr,
                                               const unsigned long long* kt, StepReport* rep) {
    const int t = threadIdx.x;
    if (t < 16) {
        rep->ctr[t] = ctr[t];
        ctr[t] = 0ull;
    } else if (t < 28) {
        // clock 6, the end of the fold, is this kernel's start
        rep->kt[t - 16] = t == 22 ? (unsigned long long)wall_clock64() : kt[t - 16];
    } else if (t == 28) {
        rep->status = *status;
    }
    __threadfence_system();
    __syncthreads();
    if (t == 0) {
        __threadfence_system();
        *(volatile unsigned int*)&rep->seq = 1u;
    }
}

// ------------------------------------------------------------ transform
// World vertices of the listed bodies.  quat_rotate (main.cpp:93-98)
// normalises q for every vertex; the normalised q is the same for all of a
// body's vertices, so it is computed once (bitwise the same doubles).
// Degeneracy of world triangles (plane_from_tri failing) only feeds the
// tri_checks counter: triangles whose body-frame plane is far from degenerate
// cannot become degenerate under a rotation plus a translation of magnitude
// <= 1e3 (rounding error bound, see build_mesh on the host), so only the
// mesh's "suspect" triangles are tested, or all of them for far-away bodies.
// One wave per body of the list (4 bodies per 256-thread workgroup): world
// vertices, AABB, and the count of non-degenerate triangles (only the mesh's
// suspect triangles can degenerate, unless the body is far from the origin).
// The xf_lane_* helpers read lane l's value (l uniform across the wave).
__device__ inline int xf_lane_i32(int x, int l) { return __builtin_amdgcn_readlane(x, l); }
__device__ inline int64_t xf_lane_i64(int64_t x, int l) {
    const uint64_t u = (uint64_t)x;
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, l);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), l);
    return (int64_t)(((uint64_t)hi << 32) | lo);
}
__device__ inline double xf_lane_f64(double x, int l) {
    return __longlong_as_double((long long)xf_lane_i64((int64_t)__double_as_longlong(x), l));
}

__global__ void __launch_bounds__(256) k_transform(DBodies B, DMeshes M, const int32_t* list, const int32_t* n_list,
                                                   unsigned long long* stamp) {
    stamp_here(stamp);
    const int64_t nb = list ? *n_list : B.n;
    const int64_t w = (int64_t)blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE;
    const int64_t nw = (int64_t)gridDim.x * (blockDim.x / WAVE);
    const int ln = lane_id();
    // the wave's bodies are it = w, w + nw, w + 2 nw, ...: lane k loads the
    // body index, mesh tables and normalised pose of the k-th of the next 64,
    // so the wave's per-body loads are the vertices alone
    for (int64_t it0 = w; it0 < nb; it0 += WAVE * nw) {
        const int64_t mine = it0 + ln * nw;
        int hb = 0, hnv = 0, hT = 0, hns = 0, hso = 0, hvo = 0, hto = 0;
        int64_t hwo = 0;
        double hp[3] = {0.0, 0.0, 0.0};
        Qt hq;
        hq.w = 1.0; hq.x = 0.0; hq.y = 0.0; hq.z = 0.0;
        if (mine < nb) {
            hb = list ? list[mine] : (int)mine;
            const int m = B.mesh[hb];
            const double* st = B.state + 16 * (int64_t)hb;
            hp[0] = st[0]; hp[1] = st[1]; hp[2] = st[2];
            Qt q; q.w = st[3]; q.x = st[4]; q.y = st[5]; q.z = st[6];
            hq = qnormalize(q);
            hnv = M.nverts[m]; hvo = M.vert_off[m]; hwo = B.wv_off[hb];
            hT = M.ntris[m]; hto = M.tri_off[m]; hns = M.nsusp[m]; hso = M.susp_off[m];
        }
        const int cnt = (int)min((int64_t)WAVE, (nb - it0 + nw - 1) / nw);
        for (int kb = 0; kb < cnt; ++kb) {
            const int b = xf_lane_i32(hb, kb);
            const V3 pos = v3(xf_lane_f64(hp[0], kb), xf_lane_f64(hp[1], kb), xf_lane_f64(hp[2], kb));
            Qt qn;
            qn.w = xf_lane_f64(hq.w, kb); qn.x = xf_lane_f64(hq.x, kb);
            qn.y = xf_lane_f64(hq.y, kb); qn.z = xf_lane_f64(hq.z, kb);
            Qt qc; qc.w = qn.w; qc.x = -qn.x; qc.y = -qn.y; qc.z = -qn.z;
            const int nv = xf_lane_i32(hnv, kb);
            const double* bv = M.verts + 3 * (int64_t)xf_lane_i32(hvo, kb);
            double* wv = B.wv + 3 * xf_lane_i64(hwo, kb);
            double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
            // TV rounds of the wave's vertices per pass: all their loads first
            constexpr int TV = 4;
            for (int v0 = ln; v0 < nv; v0 += TV * WAVE) {
                double bx[TV][3];
#pragma unroll
                for (int u = 0; u < TV; ++u) {
                    const int v = v0 + u * WAVE;
                    if (v < nv) { bx[u][0] = bv[3 * v]; bx[u][1] = bv[3 * v + 1]; bx[u][2] = bv[3 * v + 2]; }
                }
#pragma unroll
                for (int u = 0; u < TV; ++u) {
                    const int v = v0 + u * WAVE;
                    if (v < nv) {
                        Qt pv; pv.w = 0.0; pv.x = bx[u][0]; pv.y = bx[u][1]; pv.z = bx[u][2];
                        Qt r = qmul(qmul(qn, pv), qc);
                        V3 w = add(v3(r.x, r.y, r.z), pos);
                        wv[3 * v] = w.x; wv[3 * v + 1] = w.y; wv[3 * v + 2] = w.z;
                        mn[0] = fmin(mn[0], w.x); mn[1] = fmin(mn[1], w.y); mn[2] = fmin(mn[2], w.z);
                        mx[0] = fmax(mx[0], w.x); mx[1] = fmax(mx[1], w.y); mx[2] = fmax(mx[2], w.z);
                    }
                }
            }
            for (int k = 0; k < 3; ++k) {
                for (int d = 32; d >= 1; d >>= 1) {
                    mn[k] = fmin(mn[k], __shfl_xor(mn[k], d, WAVE));
                    mx[k] = fmax(mx[k], __shfl_xor(mx[k], d, WAVE));
                }
            }
            if (ln == 0) {
                double* box = B.aabb + 6 * (int64_t)b;
                for (int k = 0; k < 3; ++k) { box[k] = mn[k]; box[3 + k] = mx[k]; }
            }
            const int T = xf_lane_i32(hT, kb);
            const int32_t* ti = M.tri_idx + 3 * (int64_t)xf_lane_i32(hto, kb);
            const bool far = !(fabs(pos.x) <= 1e3 && fabs(pos.y) <= 1e3 && fabs(pos.z) <= 1e3);
            const int nchk = far ? T : xf_lane_i32(hns, kb);
            const int32_t* sl = M.susp + xf_lane_i32(hso, kb);
            int bad = 0;
            if (nchk > 0) {
                // the world vertices just written by this wave's lanes: make them
                // visible before other lanes read them back
                __threadfence_block();
                for (int k = ln; k < nchk; k += WAVE) {
                    const int t = far ? k : sl[k];
                    V3 a0 = ld3(wv + 3 * ti[3 * t]), a1 = ld3(wv + 3 * ti[3 * t + 1]), a2 = ld3(wv + 3 * ti[3 * t + 2]);
                    V3 n; double d;
                    bad += plane_of(a0, a1, a2, n, d) ? 0 : 1;
                }
                for (int d = 32; d >= 1; d >>= 1) bad += __shfl_xor(bad, d, WAVE);
            }
            if (ln == 0) B.nondeg[b] = T - bad;
        }
    }
}

// ------------------------------------------------------------ scan
// Exclusive scans of one or two int32 count arrays of the same length (*n_ptr
// or n) into out[0..n], out[n] = total.  Block sums of array a live at
// bsum + a * bstride (bstride >= blocks + 1).
__global__ void __launch_bounds__(256) k_scan_reduce(const int32_t* in0, const int32_t* in1, const int32_t* n_ptr,
                                                    int32_t n_fixed, int32_t* bsum, int32_t bstride) {
    const int n = n_ptr ? *n_ptr : n_fixed;
    const int64_t base = (int64_t)blockIdx.x * 1024;
    __shared__ int sw[2][4];
    const int na = in1 ? 2 : 1;
    for (int a = 0; a < na; ++a) {
        const int32_t* in = a ? in1 : in0;
        int s = 0;
        for (int k = 0; k < 4; ++k) {
            int64_t i = base + threadIdx.x * 4 + k;
            if (i < n) s += in[i];
        }
        for (int d = 32; d >= 1; d >>= 1) s += __shfl_xor(s, d, WAVE);
        if (lane_id() == 0) sw[a][threadIdx.x / WAVE] = s;
    }
    __syncthreads();
    if (threadIdx.x < (unsigned)na)
        bsum[threadIdx.x * (int64_t)bstride + blockIdx.x] =
            sw[threadIdx.x][0] + sw[threadIdx.x][1] + sw[threadIdx.x][2] + sw[threadIdx.x][3];
}
__global__ void __launch_bounds__(1024) k_scan_blocks(int32_t* bsum0, int32_t nb_fixed, const int32_t* n_ptr, int32_t na,
                                                      int32_t bstride) {
    // single workgroup: exclusive scan of the block sums in place; bsum[nb] = total
    __shared__ int sw[16];
    const int n = n_ptr ? *n_ptr : 0;
    const int nb = n_ptr ? (n + 1023) / 1024 : nb_fixed;
    for (int a = 0; a < na; ++a) {
        int32_t* bsum = bsum0 + a * (int64_t)bstride;
        int carry = 0;
        for (int base = 0; base < nb; base += 1024) {
            int i = base + threadIdx.x;
            int v = (i < nb) ? bsum[i] : 0;
            int x = v;
            for (int d = 1; d < WAVE; d <<= 1) {
                int y = __shfl_up(x, d, WAVE);
                if (lane_id() >= d) x += y;
            }
            if (lane_id() == WAVE - 1) sw[threadIdx.x / WAVE] = x;
            __syncthreads();
            int off = 0, tot = 0;
            for (int w = 0; w < 16; ++w) { off += (w < (int)threadIdx.x / WAVE) ? sw[w] : 0; tot += sw[w]; }
            if (i < nb) bsum[i] = carry + off + x - v;
            __syncthreads();
            carry += tot;
        }
        if (threadIdx.x == 0) bsum[nb] = carry;
    }
}
__global__ void __launch_bounds__(256) k_scan_apply(const int32_t* in0, const int32_t* in1, const int32_t* n_ptr,
                                                   int32_t n_fixed, const int32_t* bsum0, int32_t bstride,
                                                   int32_t* out0, int32_t* out1, int32_t* total_out) {
    __shared__ int sw[4];
    const int n = n_ptr ? *n_ptr : n_fixed;
    const int64_t base = (int64_t)blockIdx.x * 1024;
    if (base > n) return;
    const int nb = (n + 1023) / 1024;
    const int na = in1 ? 2 : 1;
    for (int a = 0; a < na; ++a) {
        const int32_t* in = a ? in1 : in0;
        int32_t* out = a ? out1 : out0;
        const int32_t* bsum = bsum0 + a * (int64_t)bstride;
        int v[4], s = 0;
        for (int k = 0; k < 4; ++k) {
            int64_t i = base + threadIdx.x * 4 + k;
            v[k] = (i < n) ? in[i] : 0;
            s += v[k];
        }
        int total;
        int off = block_excl<256>(s, sw, total);
        int run = bsum[blockIdx.x] + off;
        for (int k = 0; k < 4; ++k) {
            int64_t i = base + threadIdx.x * 4 + k;
            if (i < n) out[i] = run;
            run += v[k];
        }
        if (blockIdx.x == (unsigned)(nb > 0 ? nb - 1 : 0) && threadIdx.x == 0) {
            out[n] = bsum[nb];
            if (total_out && a == 0) *total_out = bsum[nb];
        }
        (void)total;
        __syncthreads();   // sw is reused by the second array
    }
}

// k_scan_apply without k_scan_blocks before it: each workgroup sums the block
// sums before its own (at most SCAN_ONE_PASS_BLOCKS of them) itself.
constexpr int SCAN_ONE_PASS_BLOCKS = 1024;
__global__ void __launch_bounds__(256) k_scan_apply_sum(const int32_t* in0, const int32_t* in1, const int32_t* n_ptr,
                                                       int32_t n_fixed, const int32_t* bsum0, int32_t bstride,
                                                       int32_t* out0, int32_t* out1, int32_t* total_out) {
    __shared__ int sw[4], sp[4];
    const int n = n_ptr ? *n_ptr : n_fixed;
    const int64_t base = (int64_t)blockIdx.x * 1024;
    if (base > n) return;
    const int nb = (n + 1023) / 1024;
    const int na = in1 ? 2 : 1;
    for (int a = 0; a < na; ++a) {
        const int32_t* in = a ? in1 : in0;
        int32_t* out = a ? out1 : out0;
        const int32_t* bsum = bsum0 + a * (int64_t)bstride;
        int pre = 0;
        for (int k = threadIdx.x; k < (int)blockIdx.x; k += 256) pre += bsum[k];
        int v[4], s = 0;
        for (int k = 0; k < 4; ++k) {
            int64_t i = base + threadIdx.x * 4 + k;
            v[k] = (i < n) ? in[i] : 0;
            s += v[k];
        }
        int before, total;
        (void)block_excl<256>(pre, sp, before);
        int run = before + block_excl<256>(s, sw, total);
        for (int k = 0; k < 4; ++k) {
            int64_t i = base + threadIdx.x * 4 + k;
            if (i < n) out[i] = run;
            run += v[k];
        }
        if (blockIdx.x == (unsigned)(nb > 0 ? nb - 1 : 0) && threadIdx.x == 0) {
            out[n] = before + total;
            if (total_out && a == 0) *total_out = before + total;
        }
        __syncthreads();   // sw and sp are reused by the second array
    }
}

void scan_i32(const int32_t* in, const int32_t* n_ptr, int32_t n_max, int32_t* tmp, int32_t* out, int32_t* total,
              hipStream_t s) {
    int nb = (n_max + 1023) / 1024;
    if (nb < 1) nb = 1;
    hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(256), 0, s, in, (const int32_t*)nullptr, n_ptr, n_max, tmp, nb + 1);
    if (nb <= SCAN_ONE_PASS_BLOCKS) {
        hipLaunchKernelGGL(k_scan_apply_sum, dim3(nb), dim3(256), 0, s, in, (const int32_t*)nullptr, n_ptr, n_max, tmp,
                           nb + 1, out, (int32_t*)nullptr, total);
        return;
    }
    hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(1024), 0, s, tmp, nb, n_ptr, 1, nb + 1);
    hipLaunchKernelGGL(k_scan_apply, dim3(nb), dim3(256), 0, s, in, (const int32_t*)nullptr, n_ptr, n_max, tmp, nb + 1,
                       out, (int32_t*)nullptr, total);
}
// two arrays of the same length n in one pass; tmp holds 2 (blocks + 1)
void scan_i32x2(const int32_t* in0, const int32_t* in1, int32_t n, int32_t* tmp, int32_t* out0, int32_t* out1,
                hipStream_t s) {
    int nb = (n + 1023) / 1024;
    if (nb < 1) nb = 1;
    hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(256), 0, s, in0, in1, (const int32_t*)nullptr, n, tmp, nb + 1);
    if (nb <= SCAN_ONE_PASS_BLOCKS) {
        hipLaunchKernelGGL(k_scan_apply_sum, dim3(nb), dim3(256), 0, s, in0, in1, (const int32_t*)nullptr, n, tmp,
                           nb + 1, out0, out1, (int32_t*)nullptr);
        return;
    }
    hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(1024), 0, s, tmp, nb, (const int32_t*)nullptr, 2, nb + 1);
    hipLaunchKernelGGL(k_scan_apply, dim3(nb), dim3(256), 0, s, in0, in1, (const int32_t*)nullptr, n, tmp, nb + 1,
                       out0, out1, (int32_t*)nullptr);
}

// ------------------------------------------------------------ broad phase

// Bounds of the particle positions and the largest radius: BP_PARTS
// workgroups reduce a strided share each; k_bp_cells combines them.
// With kt0 set it is also the step's first kernel (k_step_begin's work: the
// clock into *kt0 and the StepZero arrays zeroed, grid-stride).
__global__ void __launch_bounds__(256) k_bp_bounds_part(const double* state, const double* radius, int n,
                                                        double* part, unsigned long long* kt0, StepZero z) {
    if (kt0) {
        if (blockIdx.x == 0 && threadIdx.x == 0) *kt0 = (unsigned long long)wall_clock64();
        const int64_t t0 = blockIdx.x * (int64_t)blockDim.x + threadIdx.x, nt = (int64_t)gridDim.x * blockDim.x;
        for (int a = 0; a < 8; ++a)
            if (z.i32[a])
                for (int64_t k = t0; k < z.n32[a]; k += nt) z.i32[a][k] = 0;
        for (int a = 0; a < 2; ++a)
            if (z.u64[a])
                for (int64_t k = t0; k < z.n64[a]; k += nt) z.u64[a][k] = 0ull;
    }
    double v[7] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY, 0.0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double* p = state + 16 * (int64_t)i;
        for (int k = 0; k < 3; ++k) { v[k] = fmin(v[k], p[k]); v[3 + k] = fmax(v[3 + k], p[k]); }
        v[6] = fmax(v[6], radius[i]);
    }
    for (int d = 32; d >= 1; d >>= 1)
        for (int k = 0; k < 7; ++k) {
            const double o = __shfl_xor(v[k], d, WAVE);
            v[k] = k < 3 ? fmin(v[k], o) : fmax(v[k], o);
        }
    __shared__ double s[7][4];
    if (lane_id() == 0)
        for (int k = 0; k < 7; ++k) s[k][threadIdx.x / WAVE] = v[k];
    __syncthreads();
    if (threadIdx.x < 7) {
        const int k = threadIdx.x;
        double r = s[k][0];
        for (int w = 1; w < (int)(blockDim.x / WAVE); ++w) r = k < 3 ? fmin(r, s[k][w]) : fmax(r, s[k][w]);
        part[7 * blockIdx.x + k] = r;
    }
}

// The grid from the BP_PARTS partial bounds, by one wave (every lane ends
// with the same grid); bad: the bounds are not finite (ST_GRID).
__device__ inline DGrid bp_grid(const double* part, int nparts, int max_cells, bool& bad) {
    double v[7] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY, 0.0};
    for (int b = lane_id(); b < nparts; b += WAVE)
        for (int k = 0; k < 7; ++k) v[k] = k < 3 ? fmin(v[k], part[7 * b + k]) : fmax(v[k], part[7 * b + k]);
    for (int d = 32; d >= 1; d >>= 1)
        for (int k = 0; k < 7; ++k) {
            const double o = __shfl_xor(v[k], d, WAVE);
            v[k] = k < 3 ? fmin(v[k], o) : fmax(v[k], o);
        }
    DGrid G;
    G.nx = G.ny = G.nz = 1;
    G.ox = v[0]; G.oy = v[1]; G.oz = v[2];
    double ex = v[3] - v[0], ey = v[4] - v[1], ez = v[5] - v[2];
    bad = !(isfinite(ex) && isfinite(ey) && isfinite(ez) && isfinite(v[6]));
    // cell edge >= 2 r_max (plus margin for the rounding of (p - o) / h), so
    // every pair with |dp| <= r_i + r_j lies in adjacent cells
    double h = 2.0 * v[6] * (1.0 + 1e-6) + 1e-300;
    if (bad) { h = 1.0; ex = ey = ez = 0.0; }
    for (int it = 0; it < 64; ++it) {
        double fx = floor(ex / h) + 1.0, fy = floor(ey / h) + 1.0, fz = floor(ez / h) + 1.0;
        if (fx * fy * fz <= (double)max_cells) {
            G.nx = (int)fx; G.ny = (int)fy; G.nz = (int)fz;
            break;
        }
        h *= 1.25;
    }
    G.h = h;
    G.ncells = G.nx * G.ny * G.nz;
    return G;
}

__device__ inline void cell_coords(const DGrid& G, const double* p, int& cx, int& cy, int& cz) {
    cx = (int)floor((p[0] - G.ox) / G.h);
    cy = (int)floor((p[1] - G.oy) / G.h);
    cz = (int)floor((p[2] - G.oz) / G.h);
    cx = cx < 0 ? 0 : (cx >= G.nx ? G.nx - 1 : cx);
    cy = cy < 0 ? 0 : (cy >= G.ny ? G.ny - 1 : cy);
    cz = cz < 0 ? 0 : (cz >= G.nz ? G.nz - 1 : cz);
}

__global__ void k_fill_i32(int32_t* p, int32_t v, const int32_t* n_ptr, int64_t n) {
    int64_t m = n_ptr ? (int64_t)*n_ptr : n;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_fill_u64(unsigned long long* p, unsigned long long v, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}

// The first wave of every workgroup combines the partial bounds into the grid
// (the same doubles in every workgroup); workgroup 0 publishes it, and its
// cell count for the cell scan, for the later kernels.
__global__ void __launch_bounds__(256) k_bp_cells(BroadPhaseBufs bp) {
    __shared__ DGrid s_g;
    if (threadIdx.x < WAVE) {
        bool bad;
        const DGrid g = bp_grid(bp.part, BP_PARTS, bp.max_cells, bad);
        if (threadIdx.x == 0) {
            s_g = g;
            if (blockIdx.x == 0) {
                *bp.grid = g;
                *bp.n_adj = g.ncells;
                if (bad) atomicOr(bp.status, (int)ST_GRID);
            }
        }
    }
    __syncthreads();
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= bp.n) return;
    const DGrid G = s_g;
    int cx, cy, cz;
    cell_coords(G, bp.state + 16 * (int64_t)i, cx, cy, cz);
    int c = (cz * G.ny + cy) * G.nx + cx;
    bp.cell_of[i] = c;
    atomicAdd(&bp.cell_count[c], 1);
}

// The slot within the cell counts the cell's counter back down, which leaves
// every counter at zero for the next step's k_bp_cells (the counters are
// cleared once when allocated).  The order within a cell is free: each
// particle's partners are sorted by index in k_bp_fill.
__global__ void __launch_bounds__(256) k_bp_scatter(BroadPhaseBufs bp) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= bp.n) return;
    int c = bp.cell_of[i];
    int slot = bp.cell_start[c] + atomicSub(&bp.cell_count[c], 1) - 1;
    bp.sorted[slot] = i;
    // position and radius in cell order, for the neighbour scans
    const double* p = bp.state + 16 * (int64_t)i;
    double* o = bp.spos + 4 * (int64_t)slot;
    o[0] = p[0]; o[1] = p[1]; o[2] = p[2]; o[3] = bp.radius[i];
}

// Neighbours j of particle i (at cell-order slot si) that pass the
// reference's sphere test (main.cpp:2054-2059), evaluated for the pair
// (min(i,j), max(i,j)) on the same doubles: f(j) in cell-scan order.
template <class F>
__device__ inline void for_neighbours(const BroadPhaseBufs& bp, const DGrid& G, int i, int si, F&& f) {
    const double* qi = bp.spos + 4 * (int64_t)si;
    const double xi = qi[0], yi = qi[1], zi = qi[2], ri = qi[3];
    int cx, cy, cz;
    cell_coords(G, qi, cx, cy, cz);
    // the (up to) three cells along x of a row are consecutive cell indices,
    // so their particles are one contiguous range of the cell order: one range
    // per (dy, dz).  The nine ranges' bounds are loaded together, then each
    // range is read four particles at a time, so that a thread waits on a
    // few load rounds rather than one per candidate
    const int x0 = cx > 0 ? cx - 1 : 0, x1 = cx + 1 < G.nx ? cx + 1 : G.nx - 1;
    int s0[9], s1[9];
#pragma unroll
    for (int r = 0; r < 9; ++r) {
        const int z = cz + r / 3 - 1, y = cy + r % 3 - 1;
        const bool ok = z >= 0 && z < G.nz && y >= 0 && y < G.ny;
        const int crow = ok ? (z * G.ny + y) * G.nx : 0;
        s0[r] = ok ? bp.cell_start[crow + x0] : 0;
        s1[r] = ok ? bp.cell_start[crow + x1 + 1] : 0;
    }
    for (int r = 0; r < 9; ++r) {
        const int e = s1[r];
        for (int k0 = s0[r]; k0 < e; k0 += 4) {
            int jj[4];
            double q[4][4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                if (k0 + u < e) {
                    jj[u] = bp.sorted[k0 + u];
                    const double2* qj = reinterpret_cast<const double2*>(bp.spos + 4 * (int64_t)(k0 + u));
                    const double2 lo = qj[0], hi = qj[1];
                    q[u][0] = lo.x; q[u][1] = lo.y; q[u][2] = hi.x; q[u][3] = hi.y;
                }
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                if (k0 + u >= e) break;
                const int j = jj[u];
                if (j == i) continue;
                const bool iflow = i < j;
                const V3 pa = iflow ? v3(xi, yi, zi) : v3(q[u][0], q[u][1], q[u][2]);
                const V3 pb = iflow ? v3(q[u][0], q[u][1], q[u][2]) : v3(xi, yi, zi);
                const V3 dp = sub(pb, pa);
                const double d2 = dot(dp, dp);
                const double rs = iflow ? ri + q[u][3] : q[u][3] + ri;
                if (!(d2 > rs * rs)) f(j);
            }
        }
    }
}

constexpr int NEIGH_REG = 16;   // partners kept from the count pass and sorted in registers

// threads take particles in cell order (neighbouring threads scan the same
// cells); the first NEIGH_REG partners found are kept for k_bp_fill
// When need.list is set the same pass also builds the need lists
// (k_need_lists' test on the particle's counts and walls).
__device__ inline bool near_any_wall(const WallBufs& W, const double* st, double r);
__device__ inline void need_append(bool need, bool wall, int i, int32_t* list, int32_t* n_list, int32_t* wlist,
                                   int32_t* n_wlist);
__global__ void __launch_bounds__(256) k_bp_count(BroadPhaseBufs bp, NeedLists need) {
    const int si = blockIdx.x * blockDim.x + threadIdx.x;
    bool nd = false, wall = false;
    int i = -1;
    if (si < bp.n) {
        i = bp.sorted[si];
        const DGrid G = *bp.grid;
        int up = 0, all = 0;
        int32_t* keep = bp.nbr + (int64_t)NEIGH_REG * si;
        for_neighbours(bp, G, i, si, [&](int j) {
            if (all < NEIGH_REG) keep[all] = j;
            all++;
            if (j > i) up++;
        });
        bp.cnt_up[i] = up;
        bp.cnt_all[i] = all;
        if (need.list) {
            wall = near_any_wall(need.W, bp.state + 16 * (int64_t)i, bp.radius[i]);
            nd = all > 0 || wall;
        }
    }
    if (need.list) need_append(nd, wall, i, need.list, need.n_list, need.wlist, need.n_wlist);
}

__global__ void __launch_bounds__(128) k_bp_fill(BroadPhaseBufs bp) {
    const int si = blockIdx.x * blockDim.x + threadIdx.x;
    if (si >= bp.n) return;
    const int i = bp.sorted[si];
    const DGrid G = *bp.grid;
    if (bp.adj_off[bp.n] > bp.adj_cap || bp.pair_off[bp.n] > bp.pair_cap) {
        atomicOr(bp.status, (int)ST_PAIRS);
        return;
    }
    const int cnt = bp.cnt_all[i];
    if (cnt == 0) return;   // no partners: nothing to write
    const int ao = bp.adj_off[i];
    const int po = bp.pair_off[i];
    auto emit = [&](int k, int j, int& u) {
        bp.adj_j[ao + k] = j;
        if (j > i) {
            const int p = po + u;
            bp.pi[p] = i;
            bp.pj[p] = j;
            bp.pkey[p] = ((int64_t)i << 32) | (int64_t)(uint32_t)j;
            bp.adj_pair[ao + k] = p;
            ++u;
        } else {
            bp.adj_pair[ao + k] = -1;
        }
    };
    int u = 0;
    if (cnt <= NEIGH_REG) {
        // the partners kept by the count pass, in ascending order: insertion
        // into a sorted register array
        const int32_t* kept = bp.nbr + (int64_t)NEIGH_REG * si;
        int r[NEIGH_REG];
        int n = 0;
        for (int k = 0; k < cnt; ++k) {
            int x = kept[k];
#pragma unroll
            for (int q = 0; q < NEIGH_REG; ++q) {
                if (q < n) {
                    if (x < r[q]) { const int t = r[q]; r[q] = x; x = t; }
                } else if (q == n) {
                    r[q] = x;
                }
            }
            ++n;
        }
#pragma unroll
        for (int q = 0; q < NEIGH_REG; ++q)
            if (q < n) emit(q, r[q], u);
        return;
    }
    // more partners than the count pass kept: scan again straight into the
    // particle's adjacency slice (exactly cnt entries: the same test on the
    // same doubles), sort it in place (Shell sort, Ciura's gaps), then emit
    int32_t* js = bp.adj_j + ao;
    int n = 0;
    for_neighbours(bp, G, i, si, [&](int j) { if (n < cnt) js[n] = j; ++n; });
    constexpr int gaps[8] = {701, 301, 132, 57, 23, 10, 4, 1};
    for (int gi = 0; gi < 8; ++gi) {
        const int gp = gaps[gi];
        for (int a = gp; a < cnt; ++a) {
            const int v = js[a];
            int b = a;
            while (b >= gp && js[b - gp] > v) { js[b] = js[b - gp]; b -= gp; }
            js[b] = v;
        }
    }
    for (int k = 0; k < cnt; ++k) emit(k, js[k], u);
}

__global__ void __launch_bounds__(256) k_bp_adjfix(BroadPhaseBufs bp) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        // totals: on overflow the pair and adjacency lists were not written;
        // later stages see no pairs (the fold checks adj_off[n] against the
        // capacity) and the host grows the buffers and runs the step again
        const int np = bp.pair_off[bp.n], na = bp.adj_off[bp.n];
        const bool over = np > bp.pair_cap || na > bp.adj_cap;
        if (over) atomicOr(bp.status, (int)ST_PAIRS);
        *bp.n_pairs = over ? 0 : np;
        *bp.n_adj = over ? 0 : na;
    }
    if (i >= bp.n) return;
    if (bp.adj_off[bp.n] > bp.adj_cap || bp.pair_off[bp.n] > bp.pair_cap) return;
    const int a0 = bp.adj_off[i], a1 = bp.adj_off[i + 1];
    for (int k = a0; k < a1; ++k) {
        int j = bp.adj_j[k];
        if (j > i) continue;
        int lo = bp.pair_off[j], hi = bp.pair_off[j + 1] - 1, found = -1;
        while (lo <= hi) {
            int mid = (lo + hi) >> 1;
            int v = bp.pj[mid];
            if (v == i) { found = mid; break; }
            if (v < i) lo = mid + 1; else hi = mid - 1;
        }
        bp.adj_pair[k] = found;
    }
}


// Sphere test and triangle AABB test of the particle-wall path
// (main.cpp:2318-2323, 2338-2353): true if any wall triangle of wall w needs
// the plane section of particle i this step.
__device__ inline bool wall_tri_candidate(const double* wt9, const V3& pmn, const V3& pmx);
__device__ inline bool near_wall(const WallBufs& W, int w, const V3& pos, double r) {
    V3 dp = sub(pos, v3(W.pos[3 * w], W.pos[3 * w + 1], W.pos[3 * w + 2]));
    double d2 = dot(dp, dp);
    double rs = r + W.radius[w];
    if (d2 > rs * rs) return false;
    V3 pmn = sub(pos, v3(r, r, r)), pmx = add(pos, v3(r, r, r));
    const int t0 = W.tri_off[w], nt = W.ntri[w];
    for (int t = 0; t < nt; ++t)
        if (wall_tri_candidate(W.tris + 9 * (int64_t)(t0 + t), pmn, pmx)) return true;
    return false;
}
__device__ inline bool near_any_wall(const WallBufs& W, const double* st, double r) {
    if (W.n_walls <= 0) return false;
    const V3 pos = v3(st[0], st[1], st[2]);
    bool wall = false;
    for (int w = 0; w < W.n_walls && !wall; ++w) wall = near_wall(W, w, pos, r);
    return wall;
}

// Particles in a candidate pair or near a wall triangle (list) and those near
// a wall triangle (wlist), appended with one atomic per workgroup and list.
// The order of the lists is free: the transform and the wall kernel write
// per-body results.
// Appends i (taken by this thread when need / wall) to list / wlist: one
// atomic per list and 256-thread workgroup; every thread of the workgroup
// calls it.
__device__ inline void need_append(bool need, bool wall, int i, int32_t* list, int32_t* n_list, int32_t* wlist,
                                   int32_t* n_wlist) {
    __shared__ int s_cnt[2][4];
    __shared__ int s_base[2];
    const uint64_t bn = __ballot(need), bw = __ballot(wall);
    const int wid = threadIdx.x / WAVE, lane = lane_id();
    if (lane == 0) {
        s_cnt[0][wid] = __popcll(bn);
        s_cnt[1][wid] = __popcll(bw);
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        const int a = threadIdx.x;
        const int tot = s_cnt[a][0] + s_cnt[a][1] + s_cnt[a][2] + s_cnt[a][3];
        s_base[a] = tot > 0 ? atomicAdd(a ? n_wlist : n_list, tot) : 0;
    }
    __syncthreads();
    const uint64_t below = (1ull << lane) - 1ull;
    int on = s_base[0], ow = s_base[1];
    for (int w = 0; w < wid; ++w) {
        on += s_cnt[0][w];
        ow += s_cnt[1][w];
    }
    if (need) list[on + __popcll(bn & below)] = i;
    if (wall) wlist[ow + __popcll(bw & below)] = i;
}

__global__ void __launch_bounds__(256) k_need_lists(int32_t n, const double* state, const double* radius,
                                                    const int32_t* cnt_all, WallBufs W, int32_t* list,
                                                    int32_t* n_list, int32_t* wlist, int32_t* n_wlist,
                                                    unsigned long long* stamp) {
    stamp_here(stamp);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    bool wall = false, need = false;
    if (i < n) {
        wall = near_any_wall(W, state + 16 * (int64_t)i, radius[i]);
        need = cnt_all[i] > 0 || wall;
    }
    need_append(need, wall, i, list, n_list, wlist, n_wlist);
}

#ifdef ZDX_WALK_PROFILE
// Profiling build: device clock ticks per phase, summed over all pairs, read
// back with zdx_walk_profile().  Slots 0-7: the topology walks (slot 0: from
// the start of a pair to its first mark); slots 8-14: the narrow phase (pair
// set-up, cull A, cull B, sweep tests, exact tests, publish, pairs done).
__device__ unsigned long long g_walk_prof[16];
__device__ long long g_topo_pairs[4 * 8192];    // per topology item: S, pass A ticks, walk D ticks
__device__ long long g_walk_pairs[12 * 8192];   // walk B per pair: S, R, scratch in LDS, ticks, phases 0-7
__device__ long long g_walkd_pairs[8 * 8192];   // walk D per pair: T-junctions, adjacency, start sort, loops;
                                                // loop phase: start search, walks, Sn/Gn, force and record
// per topology item, absolute clocks: walk B's marks 0-15 (last call), chain
// start, walk D start, walk D end, chain end, hardware unit, order, components,
// snap hits
__device__ long long g_walk_abs[32 * 8192];
__device__ long long g_snap_hits[8192];
#endif

// ------------------------------------------------------------ narrow phase
#include "zdx_narrow.h"
static_assert(NQ_COUNTERS == NARROW_COUNTERS, "narrow-phase counters");

// Narrow phase, part 2 (exact tests): one thread per hit of the hit arena.
// Each hit is re-tested with the reference's exact AABB reject
// (main.cpp:2130-2132) and, when it passes, with the triangle-triangle
// intersection (main.cpp:579-631); the result goes to the hit's own slot: the
// flag, the segment in the segment arena (whose slots are the hit slots), and
// with provenance (ta, tb, endpoint codes of tri_tri_prov).
template <bool PROV>
__global__ void __launch_bounds__(256) k_narrow_exact(DBodies B, DPairs P) {
    // the used prefix of every shard of the arena, one thread per used slot:
    // the slot of linear index i is found in the prefix sums of the shards'
    // counts (after an overflow part of the arena was never written: test
    // nothing, the step is run again with a larger arena)
    static_assert(HIT_SHARDS == WAVE, "one lane per shard");
    __shared__ long long s_pre[HIT_SHARDS + 1];
    const int64_t size = P.hit_cap / P.hit_shards;
    if (threadIdx.x < WAVE) {
        const int l = threadIdx.x;
        long long c = 0;
        if (l < P.hit_shards) {
            const long long t = (long long)P.hit_top[l * HIT_TOP_STRIDE];
            c = t < size ? t : size;
        }
        long long x = c;
#pragma unroll
        for (int d = 1; d < WAVE; d <<= 1) {
            const long long y = __shfl_up(x, d, WAVE);
            if (l >= d) x += y;
        }
        s_pre[l + 1] = x;
        if (l == 0) s_pre[0] = 0;
    }
    __syncthreads();
    const int64_t lim = (*P.status & (int)ST_HITS) ? 0 : s_pre[HIT_SHARDS];
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < lim; i += (int64_t)gridDim.x * blockDim.x) {
        int lo = 0, hi = HIT_SHARDS;   // s_pre[lo] <= i < s_pre[hi]
        while (hi - lo > 1) {
            const int mid = (lo + hi) / 2;
            if (s_pre[mid] <= i) lo = mid; else hi = mid;
        }
        const int64_t h = (int64_t)lo * size + (i - s_pre[lo]);
        const int2* rec = (const int2*)(P.hits + 6 * h);
        const int2 r0 = rec[0], r1 = rec[1], r2 = rec[2];
        V3 A[3] = {ld3(B.wv + 3 * (int64_t)r0.x), ld3(B.wv + 3 * (int64_t)r0.y), ld3(B.wv + 3 * (int64_t)r1.x)};
        V3 Bt[3] = {ld3(B.wv + 3 * (int64_t)r1.y), ld3(B.wv + 3 * (int64_t)r2.x), ld3(B.wv + 3 * (int64_t)r2.y)};
        double alo[3], ahi[3], blo[3], bhi[3];
        tri_box(A, alo, ahi);
        tri_box(Bt, blo, bhi);
        // exact reference reject (main.cpp:2130-2132)
        const bool rej = ahi[0] < blo[0] || alo[0] > bhi[0] || ahi[1] < blo[1] || alo[1] > bhi[1] ||
                         ahi[2] < blo[2] || alo[2] > bhi[2];
        bool sg = false;
        V3 p0, p1;
        int c0 = 0, c1 = 0;
        if (!rej) {
            V3 nA3, nB3; double dA, dB;
            plane_of(A[0], A[1], A[2], nA3, dA);
            plane_of(Bt[0], Bt[1], Bt[2], nB3, dB);
            if (PROV) sg = tri_tri_prov(A, nA3, dA, Bt, nB3, dB, p0, p1, c0, c1);
            else sg = tri_tri(A, nA3, dA, Bt, nB3, dB, p0, p1);
        }
        P.hit_flag[h] = sg ? 1 : 0;
        if (sg) {
            double* o = P.arena + 6 * h;
            o[0] = p0.x; o[1] = p0.y; o[2] = p0.z; o[3] = p1.x; o[4] = p1.y; o[5] = p1.z;
            if (PROV) {
                const int2 t = ((const int2*)P.hit_tri)[h];
                ((int4*)P.prov)[h] = make_int4(t.x, t.y, c0, c1);
            }
        }
    }
}

// pairs with at least this many segments are walked first (their walks are
// the longest, so they should not wait for a free workgroup)
constexpr int TOPO_BIG_SEGS = 40;

// Narrow phase, part 3: one wave per pair compacts its segments in hit
// order, i.e. in the reference's (ta, tb) order, to the front of its hit
// range (in place: a segment only moves to a lower slot, and each round of 64
// reads its slots before it writes).  A pair with more than seg_cap segments
// stores none and raises ST_STAGE.  (k_topo_lists then lists the pairs with
// segments.)
template <bool PROV>
__global__ void __launch_bounds__(256) k_narrow_compact(DPairs P, int32_t seg_cap) {
    const int lane = lane_id();
    const int npairs = (*P.status & (int)ST_HITS) ? 0 : *P.n;
    const int wpb = blockDim.x / WAVE;
    for (int p = blockIdx.x * wpb + (int)threadIdx.x / WAVE; p < npairs; p += gridDim.x * wpb) {
        const int nh = P.tri_hits[p];
        if (nh == 0) continue;
        const int64_t h0 = P.hit_off[p];
        // CC chunks of WAVE hits per round: all their flags, then all their
        // segments in flight at once.  A segment only moves down (to its rank
        // within the pair), below every source of the later chunks, so the
        // in-place moves never overwrite a segment still to be read.
        constexpr int CC = 4;
        int nseg = 0;
        for (int b0 = 0; b0 < nh; b0 += CC * WAVE) {
            bool sg[CC];
#pragma unroll
            for (int u = 0; u < CC; ++u) {
                const int k = b0 + u * WAVE + lane;
                sg[u] = k < nh && P.hit_flag[h0 + k] != 0;
            }
            double sv[CC][6];
            int4 pv[CC];
#pragma unroll
            for (int u = 0; u < CC; ++u) {
                const int64_t h = h0 + b0 + u * WAVE + lane;
                pv[u] = make_int4(0, 0, 0, 0);
                if (sg[u]) {
                    for (int q = 0; q < 6; ++q) sv[u][q] = P.arena[6 * h + q];
                    if (PROV) pv[u] = ((const int4*)P.prov)[h];
                }
            }
#pragma unroll
            for (int u = 0; u < CC; ++u) {
                const unsigned long long m = __ballot(sg[u]);
                if (sg[u]) {
                    const int r = nseg + __popcll(m & ((1ull << lane) - 1ull));
                    if (r < seg_cap) {
                        double* dst = P.arena + 6 * (h0 + r);
                        for (int q = 0; q < 6; ++q) dst[q] = sv[u][q];
                        if (PROV) ((int4*)P.prov)[h0 + r] = pv[u];
                    }
                }
                nseg += __popcll(m);
            }
        }
        if (lane == 0) {
            int nstore = nseg;
            if (nstore > seg_cap) {
                atomicOr(P.status, (int)ST_STAGE);
                nstore = 0;
            }
            P.seg_off[p] = h0;
            P.nseg[p] = nstore;
        }
    }
}

// The topology's work lists: every pair with stored segments in topo_list,
// and (with P.topo_big) its index there in topo_big (TOPO_BIG_SEGS or more
// segments) or topo_small.  One thread per pair, one atomic per list and
// wave (one counter taking an atomic per pair would serialise them).
__device__ inline int wave_append(bool put, int* counter) {
    const uint64_t m = __ballot(put);
    int base = 0;
    if (lane_id() == 0 && m) base = atomicAdd(counter, __popcll(m));
    base = __shfl(base, 0, WAVE);
    return base + __popcll(m & ((1ull << lane_id()) - 1ull));
}
__global__ void __launch_bounds__(256) k_topo_lists(DPairs P) {
    const int npairs = (*P.status & (int)ST_HITS) ? 0 : *P.n;
    const int stride = gridDim.x * blockDim.x;
    for (int p0 = blockIdx.x * blockDim.x + (int)(threadIdx.x / WAVE) * WAVE; p0 < npairs; p0 += stride) {
        const int p = p0 + lane_id();
        const int ns = p < npairs ? P.nseg[p] : 0;
        const int idx = wave_append(ns > 0, P.topo_n);
        if (ns > 0) P.topo_list[idx] = p;
        if (P.topo_big) {
            const int ib = wave_append(ns >= TOPO_BIG_SEGS, P.n_big);
            const int is = wave_append(ns > 0 && ns < TOPO_BIG_SEGS, P.n_small);
            if (ns >= TOPO_BIG_SEGS) P.topo_big[ib] = idx;
            else if (ns > 0) P.topo_small[is] = idx;
        }
    }
}

// LPT order: pairs by descending estimated cost, so the persistent queues
// hand out the expensive pairs first.
// Estimated narrow-phase cost of pair p, in microseconds of one wave: its
// vertex and cull passes (per vertex, per triangle) and its float-box tests,
// with each body's kept triangles estimated as its triangles times the
// fraction of its box inside the other's (the product over the axes: the kept
// triangles are those near the overlap).  Weights from per-pair clocks of the
// one-wave kernel on MI355X; the count of tests is what varies most between
// pairs of the same meshes.  Bucket: 8 per octave, descending cost.
__device__ inline int lpt_bucket(const DPairs& P, const DMeshes& M, const DBodies& B, int p) {
    const int i = P.pi[p], j = P.pj[p];
    const int mA = B.mesh[i], mB = B.mesh[j];
    const double* a = B.aabb + 6 * (int64_t)i;
    const double* b = B.aabb + 6 * (int64_t)j;
    double fa = 1.0, fb = 1.0;
    for (int k = 0; k < 3; ++k) {
        const double ov = fmax(fmin(a[3 + k], b[3 + k]) - fmax(a[k], b[k]), 0.0);
        fa *= fmin(1.0, ov / fmax(a[3 + k] - a[k], 1e-30));
        fb *= fmin(1.0, ov / fmax(b[3 + k] - b[k], 1e-30));
    }
    const double TA = M.ntris[mA], TB = M.ntris[mB];
    const double c = 2.0 + 0.01 * (M.nverts[mA] + M.nverts[mB]) + 0.005 * (TA + TB) + 0.008 * (TA * fa) * (TB * fb);
    const int o = (int)(8.0f * __log2f((float)c));   // c >= 2: o >= 8
    return 63 - (o < 0 ? 0 : (o > 63 ? 63 : o));
}
// per-workgroup histograms first, one global atomic per (workgroup, bucket)
__global__ void __launch_bounds__(256) k_lpt_count(DPairs P, DMeshes M, DBodies B, int32_t* bucket_cnt,
                                                   unsigned long long* stamp) {
    stamp_here(stamp);
    __shared__ int h[64];
    if (threadIdx.x < 64) h[threadIdx.x] = 0;
    __syncthreads();
    const int n = *P.n;
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x)
        atomicAdd(&h[lpt_bucket(P, M, B, p)], 1);
    __syncthreads();
    if (threadIdx.x < 64 && h[threadIdx.x]) atomicAdd(&bucket_cnt[threadIdx.x], h[threadIdx.x]);
}
// order within a bucket is free (any order gives the same per-pair results)
__device__ inline void box_f(const double* lo, const double* hi, float* f);
__global__ void __launch_bounds__(256) k_lpt_scatter(DPairs P, DMeshes M, DBodies B, const int32_t* bucket_cnt,
                                                     int32_t* bucket_cur, int32_t* order, TopoBufs T,
                                                     uint8_t* cur_active, int32_t tz) {
    __shared__ int h[64], base[64], pre[64];
    const int n = *P.n;
    // the buckets' first slots: the exclusive prefix of the bucket counts (one
    // wave); bucket_cur counts each bucket's slots taken so far (from zero)
    if (threadIdx.x < 64) {
        const int c = bucket_cnt[threadIdx.x];
        int x = c;
        for (int d = 1; d < WAVE; d <<= 1) {
            const int y = __shfl_up(x, d, WAVE);
            if ((int)threadIdx.x >= d) x += y;
        }
        pre[threadIdx.x] = x - c;
    }
    __syncthreads();
    for (int p0 = blockIdx.x * blockDim.x; p0 < n; p0 += gridDim.x * blockDim.x) {
        if (threadIdx.x < 64) h[threadIdx.x] = 0;
        __syncthreads();
        const int p = p0 + threadIdx.x;
        int b = 0, r = 0;
        if (p < n) {
            b = lpt_bucket(P, M, B, p);
            r = atomicAdd(&h[b], 1);
        }
        __syncthreads();
        if (threadIdx.x < 64 && h[threadIdx.x])
            base[threadIdx.x] = pre[threadIdx.x] + atomicAdd(&bucket_cur[threadIdx.x], h[threadIdx.x]);
        __syncthreads();
        if (p < n) {
            if (tz) {
                // the topology's per-pair outputs start at zero (k_topo_zero)
                T.ncontrib[p] = 0;
                T.comps[p] = 0;
                T.loops_raw[p] = 0;
                T.tstatus[p] = 0;
                if (cur_active) cur_active[p] = 0;
            }
            const int slot = base[b] + r;
            order[slot] = p;
            if (P.items) {
                NarrowItem it;
                const int i = P.pi[p], j = P.pj[p];
                const int mA = B.mesh[i], mB = B.mesh[j];
                it.p = p; it.i = i; it.j = j;
                it.TA = M.ntris[mA]; it.TB = M.ntris[mB];
                it.VA = M.nverts[mA]; it.VB = M.nverts[mB];
                it.toA = M.tri_off[mA]; it.toB = M.tri_off[mB];
                it.woA = (int32_t)B.wv_off[i]; it.woB = (int32_t)B.wv_off[j];
                it.nondegA = B.nondeg[i];
                box_f(B.aabb + 6 * (int64_t)i, B.aabb + 6 * (int64_t)i + 3, it.boxA);
                box_f(B.aabb + 6 * (int64_t)j, B.aabb + 6 * (int64_t)j + 3, it.boxB);
                for (int k = 0; k < 8; ++k) it.pad[k] = 0;
                P.items[slot] = it;
            }
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------ topology
struct SimSink {
    double* contrib;
    int lmax;
    const Body* a;
    const Body* b;
    const Mat* ma;
    const Mat* mb;
    double dt;
    V3* ds;
    int model;
    double k_minimal;
    __device__ int operator()(int li, int, int, const V3&, const V3&, const V3& nA, const V3& xc, double area, int,
                              SArr<int>, SArr<double>, const double*) {
        Contact c = model == 1 ? pp_force_minimal(nA, area, xc, *a, *b, k_minimal)
                               : pp_force(nA, area, xc, *a, *b, *ma, *mb, dt, *ds, model == 2);
        if (li >= lmax) return TOPO_LOOP_OVERFLOW;
        double* o = contrib + 9 * (int64_t)li;
        o[0] = c.f.x; o[1] = c.f.y; o[2] = c.f.z;
        o[3] = c.r1.x; o[4] = c.r1.y; o[5] = c.r1.z;
        o[6] = c.r2.x; o[7] = c.r2.y; o[8] = c.r2.z;
        return 0;
    }
};

__device__ inline Mat load_mat(const DMat& m, int i) {
    Mat r;
    r.eqr = m.eqr[i]; r.young = m.young[i]; r.poisson = m.poisson[i]; r.mu = m.mu[i]; r.rest = m.rest[i];
    r.inv_mass = m.inv_mass[i]; r.mass = m.mass[i];
    return r;
}

// Wave-cooperative copy of n elements with four loads in flight per lane.
template <class T>
__device__ inline void wave_copy(T* dst, const T* src, int n) {
    int k = lane_id();
    for (; k + 3 * WAVE < n; k += 4 * WAVE) {
        const T a0 = src[k], a1 = src[k + WAVE], a2 = src[k + 2 * WAVE], a3 = src[k + 3 * WAVE];
        dst[k] = a0; dst[k + WAVE] = a1; dst[k + 2 * WAVE] = a2; dst[k + 3 * WAVE] = a3;
    }
    for (; k < n; k += WAVE) dst[k] = src[k];
}

// Lane policy of the walks: one wave, all lanes on the same serial path.
struct LaneDev {
    static constexpr int W = WAVE;
    __device__ int lane() const { return lane_id(); }
    __device__ uint64_t ballot(bool p) const { return __ballot(p); }
    __device__ double shfl(double v, int src) const { return __shfl(v, src, WAVE); }
    __device__ uint64_t shfl64(uint64_t v, int src) const {
        const uint32_t lo = (uint32_t)__shfl((int)(uint32_t)v, src, WAVE);
        const uint32_t hi = (uint32_t)__shfl((int)(uint32_t)(v >> 32), src, WAVE);
        return ((uint64_t)hi << 32) | lo;
    }
    __device__ int atomic_cas(int* p, int expected, int desired) const { return atomicCAS(p, expected, desired); }
    __device__ void atomic_min(int* p, int v) const { atomicMin(p, v); }
    __device__ void atomic_add(int* p, int v) const { atomicAdd(p, v); }
    // value of lane l (l the same in every lane)
    __device__ int readlane(int v, int l) const { return __builtin_amdgcn_readlane(v, l); }
    __device__ uint64_t readlane64(uint64_t v, int l) const {
        const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, l);
        const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), l);
        return ((uint64_t)hi << 32) | lo;
    }
    __device__ int scan_excl(int v, int& total) const {
        int x = v;
        for (int d = 1; d < WAVE; d <<= 1) {
            const int y = __shfl_up(x, d, WAVE);
            if (lane_id() >= d) x += y;
        }
        total = __shfl(x, WAVE - 1, WAVE);
        return x - v;
    }
    __device__ void sync() const { __syncthreads(); }
#ifdef ZDX_WALK_PROFILE
    // phase ticks accumulate in registers; the kernel stores them per pair
    mutable unsigned long long t0 = 0, t_begin = 0;
    mutable unsigned ph[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    mutable unsigned long long at[16] = {};
    __device__ void start() const {
        t0 = wall_clock64(); t_begin = t0;
        for (int k = 0; k < 8; ++k) ph[k] = 0;
        cnt = 0;
    }
    __device__ void mark(int k) const {
        asm volatile("s_waitcnt vmcnt(0) lgkmcnt(0)" ::: "memory");   // charge outstanding loads to this phase
        const unsigned long long t = wall_clock64();
        if (k < 8) ph[k] += (unsigned)(t - t0);
        if (k < 16) at[k] = t;
        t0 = t;
    }
    mutable long long cnt = 0;
    __device__ void count(int v) const { cnt += v; }
#else
    __device__ void start() const {}
    __device__ void mark(int) const {}
    __device__ void count(int) const {}
#endif
};

// ---- phased topology (zdx_topo.h): pass A, walk B, pass C, walk D.
// Pass kernels: one 256-thread workgroup per pair of the work list
// (persistent); walk kernels: one wave per pair, scratch in LDS when it fits,
// else the workgroup's slice of the global scratch.

// Snap / adjacency tolerance of a pair: mesh A's (main.cpp:2118-2121) unless
// the caller of the per-pair API gave one.
__device__ inline double topo_tol(const TopoBufs& T, const DMeshes& M, int mesh_a) {
    return T.tol_override > 0.0 ? T.tol_override : M.tol[mesh_a];
}

constexpr int TP_PASS_A_SEGS = 384;     // segments pass A stages in LDS (84 bytes each)
constexpr int TP_PASS_A_KEYS = 8192;    // LDS bytes for a unit's rows x S key matrix
constexpr int TP_PASS_A_WORDS = 256;    // snap words a unit builds in LDS at a time
constexpr int TP_PASS_A_THREADS = 1024;
constexpr int TP_PASS_A_TESTS = 4096;   // segment-pair tests per unit (sets the units of a big pair)

// Units of pass A for a pair of S segments: enough that each unit has at most
// TP_PASS_A_TESTS candidate tests (rows x S) and TP_PASS_A_WORDS snap words,
// and at least one row.
__device__ inline int tp_pa_units(int S) {
    const int64_t npt = (int64_t)S * (2 * S - 1);
    const int64_t nw = (npt + 63) >> 6;
    int64_t m = ((int64_t)S * S + TP_PASS_A_TESTS - 1) / TP_PASS_A_TESTS;
    const int64_t mw = (nw + TP_PASS_A_WORDS - 1) / TP_PASS_A_WORDS;
    if (mw > m) m = mw;
    if (m > S) m = S;
    return m < 1 ? 1 : (int)m;
}

// Pass A allocation, one thread per work-list item: the item's adjacency rows
// and snap words (bump counters top[0], top[1]) and, for a pair split into
// several units, its units 1.. in the unit list (counter top[4]).  An item that
// does not fit gets offsets -1 (the step is retried with larger buffers: the
// counters keep the totals asked for).
__global__ void __launch_bounds__(256) k_tp_alloc(DPairs P, TopoBufs T) {
    stamp_here(T.stamp_alloc);
    const int n = *P.topo_n;
    const int R = T.adj_row;
    for (int it = blockIdx.x * blockDim.x + threadIdx.x; it < n; it += gridDim.x * blockDim.x) {
        const int S = P.nseg[P.topo_list[it]];
        const long long nadj = (long long)S * (R + 1);
        const long long nw = ((long long)S * (2 * S - 1) + 63) >> 6;
        const int nu = tp_pa_units(S);
        long long ao = (long long)atomicAdd(T.top + 0, (unsigned long long)nadj);
        long long so = (long long)atomicAdd(T.top + 1, (unsigned long long)nw);
        bool ok = ao + nadj <= T.adj_cap && so + nw <= T.snap_cap;
        if (nu > 1) {
            // every slot of the reservation that lies below units_cap names this
            // item, also when the reservation straddles units_cap: pass A reads
            // every slot below units_cap, and this item's offsets are -1 then
            const long long uo = (long long)atomicAdd(T.top + 4, (unsigned long long)(nu - 1));
            if (uo + nu - 1 > T.units_cap) ok = false;
            for (int b = 1; b < nu && uo + b - 1 < T.units_cap; ++b) ((int2*)T.units)[uo + b - 1] = make_int2(it, b);
        }
        T.adj_off[it] = ok ? ao : -1;
        T.snap_off[it] = ok ? so : -1;
    }
}

// Pass A: representative-point cells, the adjacency rows (per segment the
// count and up to R accepted candidates in ascending j) and the snap hit words
// of every pair.  One 1024-thread workgroup per unit: unit 0 of every item,
// then the extra units of the big pairs.  A unit of a pair split into nu
// units takes a band of rows and a range of snap words of its own.  The
// segments and cells are staged in LDS; the candidate tests of the band run
// flat into an LDS key matrix that one thread per row then compacts; the snap
// words are built in LDS (each thread a contiguous range of endpoint pairs)
// and stored whole.
// MODE 0: every item and every extra unit; MODE 1: the items of
// P.topo_small only (fewer than TOPO_BIG_SEGS segments, so one unit each);
// MODE 2: the items of P.topo_big and every extra unit.  KEYS and WORDS size
// the LDS key matrix and snap-word chunk.
template <int NT_MAX, int KEYS, int WORDS, int MODE>
__global__ void __launch_bounds__(NT_MAX) k_tp_pass_a(DPairs P, DMeshes M, DBodies B, TopoBufs T,
                                                     int32_t lds_segs) {
    extern __shared__ char s_pa[];
    const int n = MODE == 0 ? *P.topo_n : (MODE == 1 ? *P.n_small : *P.n_big);
    const int nunits = n + (MODE == 1 ? 0 : (int)*(T.top + 4));
    const int R = T.adj_row;
    const int NT = blockDim.x;
    uint8_t* keys = (uint8_t*)(s_pa + 84 * (int64_t)lds_segs);
    unsigned long long* sw = (unsigned long long*)(s_pa + 84 * (int64_t)lds_segs + KEYS);
    for (int u = blockIdx.x; u < nunits; u += gridDim.x) {
#ifdef ZDX_WALK_PROFILE
        const unsigned long long pa_t0 = wall_clock64();
#endif
        int it = MODE == 0 ? u : (u < n ? (MODE == 1 ? P.topo_small[u] : P.topo_big[u]) : 0), b = 0;
        if (u >= n) {
            if (u - n >= T.units_cap) continue;
            const int2 e = ((const int2*)T.units)[u - n];
            it = e.x;
            b = e.y;
            // k_tp_alloc wrote every slot below units_cap this step; this only
            // guards against a slot it did not write
            if (it < 0 || it >= *P.topo_n || b < 1) continue;
        }
        const long long ao = T.adj_off[it], so = T.snap_off[it];
        if (ao < 0) continue;
        const int p = P.topo_list[it];
        const int S = P.nseg[p];
        if (b >= tp_pa_units(S)) continue;
        const int64_t sb = P.seg_off[p];
        const double h = topo_tol(T, M, B.mesh[P.pi[p]]);
        const double* gseg = P.arena + 6 * sb;
        int* gc9 = T.cell9 + 9 * sb;
        const bool staged = S <= lds_segs;
        double* seg = staged ? (double*)s_pa : (double*)gseg;
        int* c9 = staged ? (int*)(s_pa + 48 * (int64_t)lds_segs) : gc9;
        const int nu = tp_pa_units(S);
        const int r0 = (int)((int64_t)b * S / nu), r1 = (int)((int64_t)(b + 1) * S / nu);
        const int npt = S * (2 * S - 1);   // endpoint pairs of 2S points
        const int nw = (npt + 63) >> 6;
        const int w0 = (int)((int64_t)b * nw / nu), w1 = (int)((int64_t)(b + 1) * nw / nu);
        if (staged)
            for (int k = threadIdx.x; k < 6 * S; k += NT) seg[k] = gseg[k];
        __syncthreads();
        for (int k = threadIdx.x; k < S; k += NT) tp_cells(seg, k, h, c9 + 9 * k);
        __syncthreads();
        int* cntp = T.adj + ao;
        int* rows = cntp + S;
        const int nr = r1 - r0;
        if (nr * S <= KEYS) {
            for (int k = threadIdx.x; k < nr * S; k += NT) {
                const int i = r0 + k / S, j = k - (k / S) * S;
                keys[k] = (uint8_t)tp_adj_key(seg, c9, h, i, j);
            }
            __syncthreads();
            for (int i = r0 + (int)threadIdx.x; i < r1; i += NT) {
                const uint8_t* kr = keys + (i - r0) * S;
                int* out = rows + (int64_t)i * R;
                int m = 0;
                for (int j = 0; j < S; ++j) {
                    const int key = kr[j];
                    if (key == TP_NOADJ) continue;
                    if (m < R) out[m] = j | (key << 24);
                    ++m;
                }
                cntp[i] = m;
                if (m > R) atomicMax(T.top + 3, (unsigned long long)m);
            }
        } else {
            for (int i = r0 + (int)threadIdx.x; i < r1; i += NT) {
                const int m = tp_adj_row(seg, c9, h, S, i, rows + (int64_t)i * R, R);
                cntp[i] = m;
                if (m > R) atomicMax(T.top + 3, (unsigned long long)m);
            }
        }
        // snap hits of the unit's words, WORDS at a time: each
        // thread a contiguous range of pair numbers t (a(a-1)/2 + b) sets bits
        // in LDS, then the words are stored whole
        uint64_t* w = T.snapw + so;
        for (int c0 = w0; c0 < w1; c0 += WORDS) {
            const int c1 = c0 + WORDS < w1 ? c0 + WORDS : w1;
            for (int k = threadIdx.x; k < c1 - c0; k += NT) sw[k] = 0ull;
            __syncthreads();
            const int tb = 64 * c0, te = 64 * c1 < npt ? 64 * c1 : npt;
            const int per = (te - tb + NT - 1) / NT;
            int t = tb + (int)threadIdx.x * per;
            const int t1 = t + per < te ? t + per : te;
            if (t < t1) {
                int a, bb;
                tp_pair_of(t, a, bb);
                for (; t < t1; ++t) {
                    if (tp_snap_hit(seg, c9, h, a, bb)) atomicOr(&sw[(t >> 6) - c0], 1ull << (t & 63));
                    if (++bb == a) { ++a; bb = 0; }
                }
            }
            __syncthreads();
            for (int k = threadIdx.x; k < c1 - c0; k += NT) w[c0 + k] = sw[k];
            __syncthreads();
        }
#ifdef ZDX_WALK_PROFILE
        if (threadIdx.x == 0 && it < 8192 && b == 0) {
            g_topo_pairs[4 * it] = S;
            g_topo_pairs[4 * it + 1] = (long long)(wall_clock64() - pa_t0);
        }
#endif
    }
}

// Snap clusters of the big-list pairs, before their walks: one 1024-thread
// workgroup per pair takes all of the pair's snap hits (pass A's words) and
// links them lock-free in LDS (a root only ever points to a smaller index, so
// every cluster ends at its smallest point), regardless of the pair's
// components; walk B uses the result when no cluster spans two components.
constexpr int CC_THREADS = 512, CC_POINTS = 8192;
__global__ void __launch_bounds__(CC_THREADS) k_tp_snap_cc(DPairs P, TopoBufs T) {
    __shared__ int s_lab[CC_POINTS];
    __shared__ int s_hl[CC_THREADS];
    __shared__ int s_wave[CC_THREADS / WAVE];
    const int nb = *P.n_big;
    const int tid = threadIdx.x;
    auto root = [&](int x) {   // with pointer jumping (see tp_components_and_vertices)
        int y = s_lab[x];
        if (y == x) return x;
        int z = s_lab[y];
        while (z != y) {
            s_lab[x] = z;
            x = y; y = z; z = s_lab[y];
        }
        return y;
    };
    for (int u = blockIdx.x; u < nb; u += gridDim.x) {
        const int it = P.topo_big[u];
        const int p = P.topo_list[it];
        const int n = 2 * P.nseg[p];
        int* glab = T.lab + 2 * P.seg_off[p];
        const long long ao = T.adj_off[it], so = T.snap_off[it];
        if (ao < 0 || n > CC_POINTS) {
            if (tid == 0) glab[0] = -1;
            continue;
        }
        for (int k = tid; k < n; k += CC_THREADS) s_lab[k] = k;
        __syncthreads();
        const int64_t nwords = ((int64_t)n * (n - 1) / 2 + 63) >> 6;
        const uint64_t* w = T.snapw + so;
        for (int64_t wb = 0; wb < nwords; wb += CC_THREADS) {
            uint64_t word = wb + tid < nwords ? w[wb + tid] : 0ull;
            int total;
            int g = block_excl<CC_THREADS>(__popcll(word), s_wave, total);
            for (int j0 = 0; j0 < total; j0 += CC_THREADS) {
                while (word && g < j0 + CC_THREADS) {
                    s_hl[g - j0] = (tid << 6) + __builtin_ctzll(word);
                    word &= word - 1;
                    ++g;
                }
                __syncthreads();
                if (j0 + tid < total) {
                    int a, b;
                    tp_pair_of((wb << 6) + s_hl[tid], a, b);
                    int ra = root(a), rb = root(b);
                    while (ra != rb) {
                        if (ra < rb) { const int x = ra; ra = rb; rb = x; }
                        const int old = atomicCAS(&s_lab[ra], ra, rb);
                        if (old == ra) break;
                        ra = root(old);
                        rb = root(rb);
                    }
                }
                __syncthreads();
            }
        }
        for (int k = tid; k < n; k += CC_THREADS) glab[k] = root(k);
        __syncthreads();
    }
}

// Sensitivity runs: Jacobian of every segment endpoint (endpoint_jacobian)
// from the narrow phase's provenance; body A of a pair is pi, B is pj.
__global__ void __launch_bounds__(256) k_tp_endpoint_jac(DPairs P, DMeshes M, DBodies B, TopoBufs T) {
    const int n = *P.topo_n;
    for (int it = blockIdx.x; it < n; it += gridDim.x) {
        const int p = P.topo_list[it];
        const int i = P.pi[p], j = P.pj[p];
        const int S = P.nseg[p];
        const int64_t sb = P.seg_off[p];
        const int mA = B.mesh[i], mB = B.mesh[j];
        const int32_t* tiA = M.tri_idx + 3 * (int64_t)M.tri_off[mA];
        const int32_t* tiB = M.tri_idx + 3 * (int64_t)M.tri_off[mB];
        const double* wvA = B.wv + 3 * B.wv_off[i];
        const double* wvB = B.wv + 3 * B.wv_off[j];
        const V3 cA = ld3(B.state + 16 * (int64_t)i), cB = ld3(B.state + 16 * (int64_t)j);
        for (int k = threadIdx.x; k < 2 * S; k += blockDim.x) {
            const int s = k >> 1, e = k & 1;
            const int32_t* pv = P.prov + 4 * (sb + s);
            V3 A[3], Bt[3];
            load_tri(wvA, tiA, pv[0], A);
            load_tri(wvB, tiB, pv[1], Bt);
            V3 nA, nB;
            double dA, dB;
            plane_of(A[0], A[1], A[2], nA, dA);
            plane_of(Bt[0], Bt[1], Bt[2], nB, dB);
            const V3 q = ld3(P.arena + 6 * (sb + s) + 3 * e);
            endpoint_jacobian(pv[2 + e], A, Bt, nA, nB, q, cA, cB, T.ej + 72 * (sb + s) + 36 * e);
        }
    }
}

// Walk scratch of one pair: LDS when it fits, else this workgroup's global
// slice.  The two cases call the walk separately so each inlined copy
// addresses its scratch directly (LDS instructions for the first).
template <class Body>
__device__ inline auto with_walk_scratch(const TopoBufs& T, int64_t ni, int64_t nd, char* lds, int lds_bytes,
                                         Body&& body) {
    const int64_t ib = (ni * 4 + 7) & ~(int64_t)7;
    if (ib + nd * 8 <= lds_bytes) return body((int*)lds, (double*)(lds + ib));
    return body(T.scr_i + (int64_t)blockIdx.x * T.scr_ni, T.scr_d + (int64_t)blockIdx.x * T.scr_nd);
}

struct WalkB {
    int ncomp;       // -1: not evaluated (the step is retried with larger buffers)
    long long to;    // first T-junction word of the item, -1: did not fit
};
// pre: the pair's snap clusters were found by k_tp_snap_cc; SENS = false
// compiles out the endpoint sensitivities (the time step never asks for them)
template <bool SENS = true>
__device__ inline WalkB tp_walk_b_item(const DPairs& P, const DMeshes& M, const DBodies& B, const TopoBufs& T,
                                       int32_t split, int32_t lds_bytes, char* s_walk, const LaneDev& wp, int it,
                                       bool pre = false) {
    const int R = T.adj_row;
    const int p = P.topo_list[it];
    const int S = P.nseg[p];
    const int64_t sb = P.seg_off[p];
    const double h = topo_tol(T, M, B.mesh[P.pi[p]]);
    const long long ao = T.adj_off[it], so = T.snap_off[it];
    int ncomp = -1, tjn = 0;
    wp.start();
    if (ao >= 0 && S <= T.seg_max) {
        const double* seg = P.arena + 6 * sb;
        ncomp = with_walk_scratch(T, tp_walk_b_ints(S, R), tp_walk_b_dbls(S), s_walk, lds_bytes,
                                  [&](int* I, double*) {
            int* adj = I + tp_b_ints(S);   // counts (S) then rows (S x R)
            const int* src = T.adj + ao;
            wave_copy(adj, src, S * (R + 1));
            __syncthreads();
            wp.mark(0);
            int ok = 1;
            for (int k = lane_id(); k < S; k += WAVE) ok &= adj[k] <= R;
            if (!__all(ok)) return -1;
            return tp_components_and_vertices(wp, seg, S, h, split != 0, adj, adj + S, R, T.snapw + so, I,
                                              T.comp + TPC_N * sb, T.V + 6 * sb, T.vcell + 6 * sb,
                                              T.E + 2 * sb, &tjn, SENS && T.ej ? T.ej + 72 * sb : nullptr,
                                              SENS && T.vj ? T.vj + 72 * sb : nullptr, pre ? T.lab + 2 * sb : nullptr);
        });
    }
#ifdef ZDX_WALK_PROFILE
    if (lane_id() == 0 && it < 8192) {
        const int64_t ib = ((int64_t)tp_walk_b_ints(S, R) * 4 + 7) & ~(int64_t)7;
        g_walk_pairs[12 * it] = S;
        g_walk_pairs[12 * it + 1] = R;
        g_walk_pairs[12 * it + 2] = ib + tp_walk_b_dbls(S) * 8 <= lds_bytes ? 1 : 0;
        g_walk_pairs[12 * it + 3] = (long long)(wall_clock64() - wp.t_begin);
        for (int k = 0; k < 8; ++k) g_walk_pairs[12 * it + 4 + k] = wp.ph[k];
        for (int k = 0; k < 16; ++k) g_walk_abs[32 * it + k] = (long long)wp.at[k];
    }
    if (ao >= 0 && it < 8192) {
        const int64_t np2 = (int64_t)(2 * S) * (2 * S - 1) / 2, nw = (np2 + 63) >> 6;
        long long c = 0;
        for (int64_t w = lane_id(); w < nw; w += WAVE) c += __popcll(T.snapw[so + w]);
        for (int d = 32; d; d >>= 1) c += __shfl_xor(c, d);
        if (lane_id() == 0) g_snap_hits[it] = c;
    }
#endif
    long long to = -1;
    if (lane_id() == 0) {
        if (ncomp >= 0) {
            to = (long long)atomicAdd(T.top + 2, (unsigned long long)tjn);
            if (to + tjn > T.tj_cap) to = -1;
        }
        T.ncomp[it] = ncomp;
        T.tj_off[it] = ncomp >= 0 ? to : -1;
    }
    to = __shfl(to, 0, WAVE);
    __syncthreads();
    wp.mark(4);
    return WalkB{ncomp, ncomp >= 0 ? to : -1};
}

// Walk B: components, snapped vertices and edges of each pair; allocates the
// pair's T-junction words.  The adjacency rows are copied into the scratch.
__global__ void __launch_bounds__(64) k_tp_walk_b(DPairs P, DMeshes M, DBodies B, TopoBufs T, int32_t split,
                                                  int32_t lds_bytes) {
    extern __shared__ char s_walk[];
    const int n = *P.topo_n;
    LaneDev wp;
    for (int it = blockIdx.x; it < n; it += gridDim.x) tp_walk_b_item(P, M, B, T, split, lds_bytes, s_walk, wp, it);
}

// Pass C: T-junction hit words of every component of an item; each wave of
// the workgroup covers 64 vertices of one edge row per ballot.
// lds / lds_bytes (optional): a component's vertices, cells and edges are
// staged there when they fit (the walks' scratch is free between walk B and
// walk D), so the ne x nv tests read LDS instead of global memory.
__device__ inline void tp_pass_c_item(const DPairs& P, const DMeshes& M, const DBodies& B, const TopoBufs& T, int it,
                                      int ncomp, long long to, char* lds = nullptr, int lds_bytes = 0) {
    if (ncomp <= 0 || to < 0) return;
    const int p = P.topo_list[it];
    const int64_t sb = P.seg_off[p];
    const double h = topo_tol(T, M, B.mesh[P.pi[p]]);
    const int* comp = T.comp + TPC_N * sb;
    for (int q = 0; q < ncomp; ++q) {
        const int* cr = comp + TPC_N * q;
        const int nv = cr[TPC_NV], ne = cr[TPC_NE];
        const double* V = T.V + 6 * sb + 3 * (int64_t)cr[TPC_VB];
        const int* vc = T.vcell + 6 * sb + 3 * (int64_t)cr[TPC_VB];
        const int* E = T.E + 2 * sb + 2 * (int64_t)cr[TPC_EB];
        uint64_t* w = T.tjw + to + cr[TPC_TW];
        const int rl = 64 * tp_row_words(nv);
        const int items = ne * rl;
        if (lds && 24 * nv + 12 * nv + 8 * ne <= lds_bytes) {
            // the component in LDS; one thread per (edge, word of 64 vertices):
            // the edge's cell range once, then its vertices in order, their
            // cells read by every thread at once
            double* lv = (double*)lds;
            int* lc = (int*)(lds + 24 * nv);
            int* le = lc + 3 * nv;
            for (int k = threadIdx.x; k < 3 * nv; k += blockDim.x) { lv[k] = V[k]; lc[k] = vc[k]; }
            for (int k = threadIdx.x; k < 2 * ne; k += blockDim.x) le[k] = E[k];
            __syncthreads();
            const int rw = rl >> 6;
            for (int u = threadIdx.x; u < ne * rw; u += blockDim.x) {
                const int e = u / rw, wi = u - e * rw;
                int cb[6];
                tp_tj_box(lv, le, h, e, cb);
                const int a = le[2 * e], b = le[2 * e + 1];
                const int v0 = 64 * wi, v1 = nv < v0 + 64 ? nv : v0 + 64;
                // cell test of four vertices at a time (their loads in flight
                // together); the distance test for those that pass
                uint64_t cand = 0;
                for (int vb = v0; vb < v1; vb += 4) {
                    int c[12];
                    for (int r = 0; r < 4; ++r) {
                        const int v = vb + r < v1 ? vb + r : vb;
                        c[3 * r] = lc[3 * v]; c[3 * r + 1] = lc[3 * v + 1]; c[3 * r + 2] = lc[3 * v + 2];
                    }
                    for (int r = 0; r < 4; ++r) {
                        const int v = vb + r;
                        const bool in = v < v1 && v != a && v != b && c[3 * r] >= cb[0] && c[3 * r] <= cb[3] &&
                                        c[3 * r + 1] >= cb[1] && c[3 * r + 1] <= cb[4] && c[3 * r + 2] >= cb[2] &&
                                        c[3 * r + 2] <= cb[5];
                        if (in) cand |= 1ull << (v - v0);
                    }
                }
                uint64_t word = 0;
                for (uint64_t m = cand; m; m &= m - 1) {
                    const int v = v0 + (int)__builtin_ctzll(m);
                    if (tp_tj_near(lv, h, a, b, v)) word |= 1ull << (v - v0);
                }
                w[u] = word;
            }
        } else {
            for (int kb = 0; kb < items; kb += blockDim.x) {
                const int t = kb + threadIdx.x;
                const int e = t / rl, v = t - e * rl;
                const bool hit = t < items && v < nv && tp_tj_hit(V, vc, E, h, e, v);
                const uint64_t m = __ballot(hit);
                if (lane_id() == 0 && t < items) w[t >> 6] = m;
            }
        }
        __syncthreads();   // the staged component is replaced by the next one
    }
}
__global__ void __launch_bounds__(256) k_tp_pass_c(DPairs P, DMeshes M, DBodies B, TopoBufs T) {
    const int n = *P.topo_n;
    for (int it = blockIdx.x; it < n; it += gridDim.x) tp_pass_c_item(P, M, B, T, it, T.ncomp[it], T.tj_off[it]);
}

// Walk D with its inputs copied to the scratch: edges, vertex cells.
template <class Sink, bool SENS = true>
__device__ inline TopoOut tp_walk_d(const LaneDev& wp, const TopoBufs& T, int S, int ncomp, int64_t sb, long long to,
                                    double h, double area_eps, const V3& dir, char* lds, int lds_bytes, Sink& sink) {
    return with_walk_scratch(T, tp_walk_d_ints(S), tp_walk_d_dbls(S), lds, lds_bytes, [&](int* I, double* D) {
        int* El = I + tp_d_ints(S);     // 2S
        int* vcl = El + 2 * S;          // 6S
        const int* comp = T.comp + TPC_N * sb;
        const int ne_all = comp[TPC_N * (ncomp - 1) + TPC_EB] + comp[TPC_N * (ncomp - 1) + TPC_NE];
        const int nv_all = comp[TPC_N * (ncomp - 1) + TPC_VB] + comp[TPC_N * (ncomp - 1) + TPC_NV];
        wave_copy(El, T.E + 2 * sb, 2 * ne_all);
        wave_copy(vcl, T.vcell + 6 * sb, 3 * nv_all);
        __syncthreads();
        return tp_loops(wp, S, ncomp, h, area_eps, dir, comp, T.V + 6 * sb, vcl, El, T.tjw + to, I, D, sink,
                        SENS && T.vj ? T.vj + 72 * sb : nullptr, D + tp_d_dbls(S));
    });
}

__device__ inline void tp_walk_d_sim_item(const DPairs& P, const DMeshes& M, const DBodies& B, const DMat& mat,
                                          const HistBufs& H, const TopoBufs& T, const SimParams& sp, int nprev,
                                          int32_t lds_bytes, char* s_walk, const LaneDev& wp, int it, int ncomp,
                                          long long to) {
    const int p = P.topo_list[it];
    const int i = P.pi[p], j = P.pj[p];
    const int S = P.nseg[p];
    const int mA = B.mesh[i];
    const Body a = load_body(B.state + 16 * (int64_t)i);
    const Body b = load_body(B.state + 16 * (int64_t)j);
    const Mat ma = load_mat(mat, i), mb = load_mat(mat, j);
    // previous tangential spring of this pair (the reference's map entry,
    // erased when the pair was not in contact on the previous step)
    const int64_t key = ((int64_t)i << 32) | (int64_t)(uint32_t)j;
    V3 ds = v3(0.0, 0.0, 0.0);
    {
        // 64-way search of the sorted keys: each round the lanes read 64
        // evenly spaced keys of [lo, hi) and keep the slice that can hold key
        int lo = 0, hi = nprev;
        while (hi - lo > WAVE) {
            const int64_t span = hi - lo;
            const int pos = lo + (int)(span * lane_id() / WAVE);
            const uint64_t le = __ballot(H.prev_key[pos] <= key);   // lanes in ascending key order
            if (le == 0ull) { hi = lo; break; }
            const int l = 63 - __builtin_clzll(le);                   // last lane with a key <= key
            const int nlo = lo + (int)(span * l / WAVE);
            hi = l == WAVE - 1 ? hi : lo + (int)(span * (l + 1) / WAVE);
            lo = nlo;
        }
        const int pos = lo + lane_id();
        const bool hit = pos < hi && H.prev_key[pos] == key;
        const uint64_t hm = __ballot(hit);
        if (hm) {
            const int mid = lo + __builtin_ctzll(hm);
            if (H.prev_active[mid]) ds = ld3(H.prev_ds + 3 * (int64_t)mid);
        }
    }
    SimSink sink{T.contrib + 9 * (int64_t)p * T.lmax, T.lmax, &a, &b, &ma, &mb, sp.dt, &ds, sp.force_model,
                 sp.k_minimal};
    wp.start();
    TopoOut o;
    o.comps = 0; o.loops_raw = 0; o.loops = 0; o.status = 0;
    if (ncomp < 0 || to < 0) o.status = TOPO_E2_OVERFLOW;
    else if (ncomp > 0)
        o = tp_walk_d<SimSink, false>(wp, T, S, ncomp, P.seg_off[p], to, topo_tol(T, M, mA), M.area_eps[mA],
                                      sub(b.pos, a.pos), s_walk,
                      lds_bytes, sink);
#ifdef ZDX_WALK_PROFILE
    if (lane_id() == 0 && it < 8192) {
        g_topo_pairs[4 * it + 2] = (long long)(wall_clock64() - wp.t_begin);
        g_walk_abs[32 * it + 17] = (long long)wp.t_begin;
        g_walk_abs[32 * it + 18] = (long long)wall_clock64();
        for (int k = 0; k < 8; ++k) g_walkd_pairs[8 * it + k] = wp.ph[k];
        g_walk_abs[32 * it + 23] = o.loops;
        g_walk_abs[32 * it + 24] = o.loops_raw;
        g_walk_abs[32 * it + 25] = wp.cnt;
    }
#endif
    if (lane_id() == 0) {
        T.ncontrib[p] = o.loops < T.lmax ? o.loops : T.lmax;
        T.comps[p] = o.comps;
        T.loops_raw[p] = o.loops_raw;
        T.tstatus[p] = o.status;
        H.cur_ds[3 * (int64_t)p] = ds.x; H.cur_ds[3 * (int64_t)p + 1] = ds.y; H.cur_ds[3 * (int64_t)p + 2] = ds.z;
        H.cur_active[p] = o.loops > 0 ? 1 : 0;
        if (o.status) atomicOr(P.status, (int)ST_TOPO);
    }
    __syncthreads();
}
__global__ void __launch_bounds__(64) k_tp_walk_d_sim(DPairs P, DMeshes M, DBodies B, DMat mat, HistBufs H,
                                                      TopoBufs T, SimParams sp, int32_t lds_bytes) {
    extern __shared__ char s_walk[];
    const int n = *P.topo_n;
    const int nprev = *H.prev_n;
    LaneDev wp;
    for (int it = blockIdx.x; it < n; it += gridDim.x)
        tp_walk_d_sim_item(P, M, B, mat, H, T, sp, nprev, lds_bytes, s_walk, wp, it, T.ncomp[it], T.tj_off[it]);
}

// Walk B, pass C and walk D of one item in a row on one wave (the time step):
// each pair goes through its whole chain without waiting for the slowest
// pair of each phase.  The walks' LDS scratch is reused from one to the next.
// mode 0: every item, the big list's first; 1: the small list; 2: the big list.
// WPS: waves per SIMD the registers are sized for -- 1 lets a long chain use
// all it needs (over 256 VGPRs); 2 fits twice as many of the short chains of
// the small list at once, each somewhat slower.
template <int WPS>
__global__ void __launch_bounds__(64, WPS) k_tp_walk_bcd_sim(DPairs P, DMeshes M, DBodies B, DMat mat, HistBufs H,
                                                             TopoBufs T, SimParams sp, int32_t lds_bytes, int32_t mode) {
    extern __shared__ char s_walk[];
    const int nprev = *H.prev_n;
    const int nbig = P.topo_big ? *P.n_big : 0;
    const int n = mode == 0 ? *P.topo_n : (mode == 1 ? *P.n_small : nbig);
    LaneDev wp;
    for (int k = blockIdx.x; k < n; k += gridDim.x) {
        const bool big = mode == 2 || (mode == 0 && k < nbig);
        const int it = mode == 1 ? P.topo_small[k]
                                 : (!P.topo_big ? k : (k < nbig ? P.topo_big[k] : P.topo_small[k - nbig]));
#ifdef ZDX_WALK_PROFILE
        const unsigned long long chain_t0 = wall_clock64();
#endif
        const WalkB r = tp_walk_b_item<false>(P, M, B, T, sp.split, lds_bytes, s_walk, wp, it, big && P.topo_big);
        __syncthreads();
        tp_pass_c_item(P, M, B, T, it, r.ncomp, r.to, s_walk, lds_bytes);
        __syncthreads();
        tp_walk_d_sim_item(P, M, B, mat, H, T, sp, nprev, lds_bytes, s_walk, wp, it, r.ncomp, r.to);
#ifdef ZDX_WALK_PROFILE
        if (lane_id() == 0 && it < 8192) {
            g_topo_pairs[4 * it + 3] = (long long)(wall_clock64() - chain_t0);
            g_walk_abs[32 * it + 16] = (long long)chain_t0;
            g_walk_abs[32 * it + 19] = (long long)wall_clock64();
            g_walk_abs[32 * it + 20] = (long long)__smid();
            g_walk_abs[32 * it + 21] = k;
            g_walk_abs[32 * it + 22] = r.ncomp;
        }
#endif
    }
}

// Outputs of pairs the topology kernels do not visit (no segments): no
// loops, no history entry (the reference's map entry is only created for an
// accepted loop and erased otherwise).
__global__ void k_topo_zero(DPairs P, TopoBufs T, uint8_t* cur_active, unsigned long long* stamp) {
    stamp_here(stamp);
    const int n = *P.n;
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
        T.ncontrib[p] = 0;
        T.comps[p] = 0;
        T.loops_raw[p] = 0;
        T.tstatus[p] = 0;
        if (cur_active) cur_active[p] = 0;
    }
}

struct ApiSink {
    double* rec;
    double* vtx;
    int lmax, vmax, nv;
    V3 posA, posB;
    double k;
    double* ljac;   // lmax * 72 of this pair, or nullptr
    __device__ int operator()(int li, int comp, int raw, const V3& Sn, const V3& Gn, const V3& nA, const V3& xc,
                              double area, int plen, SArr<int> path, SArr<double> Vs, const double* jac) {
        if (li >= lmax) return TOPO_LOOP_OVERFLOW;
        if (ljac && jac)
            for (int c = lane_id(); c < 72; c += WAVE) ljac[72 * (int64_t)li + c] = jac[c];
        double* r = rec + ZDX_LOOP_REC_D * (int64_t)li;
        r[0] = Sn.x; r[1] = Sn.y; r[2] = Sn.z; r[3] = Gn.x; r[4] = Gn.y; r[5] = Gn.z;
        r[6] = nA.x; r[7] = nA.y; r[8] = nA.z; r[9] = xc.x; r[10] = xc.y; r[11] = xc.z;
        r[12] = area;
        // minimal-energy force, polyline.py:676-679: Fn = 0.5 k |Sn|, along n
        double Fn = 0.5 * k * area;
        V3 f = mul(nA, Fn);
        V3 rA = sub(xc, posA), rB = sub(xc, posB);
        V3 tA = cross(rA, mul(f, -1.0)), tB = cross(rB, f);
        r[13] = Fn; r[14] = f.x; r[15] = f.y; r[16] = f.z;
        r[17] = tA.x; r[18] = tA.y; r[19] = tA.z; r[20] = tB.x; r[21] = tB.y; r[22] = tB.z;
        r[23] = comp; r[24] = plen; r[25] = raw; r[26] = 0.0; r[27] = 0.0;
        if (vtx) {
            if (nv + plen > vmax) return TOPO_VTX_OVERFLOW;
            for (int m = 0; m < plen; ++m) {
                int v = path[m];
                double* o = vtx + 3 * (int64_t)(nv + m);
                o[0] = Vs[3 * v]; o[1] = Vs[3 * v + 1]; o[2] = Vs[3 * v + 2];
            }
            nv += plen;
        }
        return 0;
    }
    static constexpr int ZDX_LOOP_REC_D = 28;
};

__global__ void __launch_bounds__(64) k_tp_walk_d_api(DPairs P, DMeshes M, DBodies B, TopoBufs T, int32_t lds_bytes) {
    extern __shared__ char s_walk[];
    const int n = *P.topo_n;
    LaneDev wp;
    for (int it = blockIdx.x; it < n; it += gridDim.x) {
        const int p = P.topo_list[it];
        const int i = P.pi[p], j = P.pj[p];
        const int S = P.nseg[p];
        const int mA = B.mesh[i];
        const V3 pa = ld3(B.state + 16 * (int64_t)i), pb = ld3(B.state + 16 * (int64_t)j);
        ApiSink sink{T.rec + 28 * (int64_t)p * T.lmax, T.vtx ? T.vtx + 3 * (int64_t)p * T.vmax : nullptr,
                     T.lmax, T.vmax, 0, pa, pb, T.k_stiff, T.ljac ? T.ljac + 72 * (int64_t)p * T.lmax : nullptr};
        const int ncomp = T.ncomp[it];
        const long long to = T.tj_off[it];
        TopoOut o;
        o.comps = 0; o.loops_raw = 0; o.loops = 0; o.status = 0;
        if (ncomp < 0 || to < 0) o.status = TOPO_E2_OVERFLOW;
        else if (ncomp > 0)
            o = tp_walk_d(wp, T, S, ncomp, P.seg_off[p], to, topo_tol(T, M, mA), M.area_eps[mA], sub(pb, pa), s_walk,
                          lds_bytes, sink);
        if (lane_id() == 0) {
            T.ncontrib[p] = o.loops;
            T.comps[p] = o.comps;
            T.loops_raw[p] = o.loops_raw;
            T.tstatus[p] = o.status;
            if (o.status) atomicOr(P.status, (int)ST_TOPO);
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------ walls
__device__ inline bool wall_tri_candidate(const double* wt9, const V3& pmn, const V3& pmx) {
    V3 wt[3] = {ld3(wt9), ld3(wt9 + 3), ld3(wt9 + 6)};
    double lo[3], hi[3];
    tri_box(wt, lo, hi);
    return !(pmx.x < lo[0] || pmn.x > hi[0] || pmx.y < lo[1] || pmn.y > hi[1] || pmx.z < lo[2] || pmn.z > hi[2]);
}

// One wave per listed particle (a particle near at least one wall triangle).
// Entries of particles not in the list were zeroed before this kernel.
// Registers capped at 96 (five waves per SIMD): the kernel runs beside the
// narrow phase on the second stream, and at its natural 206 VGPRs a walls wave
// left room for only two of the one-wave narrow kernel's waves on its SIMD
// instead of four (the narrow phase is 3 % faster this way; the walls, which
// end well before it, spill a little).
__global__ void __launch_bounds__(64, 5) k_walls(DBodies B, DMeshes M, DMat mat, WallBufs W, HistBufs H, double* wcontrib,
                                              int32_t* wnc, const int32_t* list, const int32_t* n_list, double* scr,
                                              int32_t pts_cap, SimParams sp, int32_t* status) {
    LaneDev wp;
    const int nl = *n_list;
    const int64_t per = 3 * (int64_t)pts_cap + 4 * (int64_t)(pts_cap + 8);
    SArr<double> PTS; PTS.p = scr + (int64_t)blockIdx.x * per; PTS.s = 1;
    SArr<double> PA = PTS.at(3 * (int64_t)pts_cap);
    SArr<double> PB = PA.at(2 * (int64_t)(pts_cap + 8));
    for (int it = blockIdx.x; it < nl; it += gridDim.x) {
        const int i = list[it];
        Body p = load_body(B.state + 16 * (int64_t)i);
        const double r = mat.radius[i];
        const int m = B.mesh[i];
        const double tol = M.tol[m];
        const double* wv = B.wv + 3 * B.wv_off[i];
        const int32_t* ti = M.tri_idx + 3 * (int64_t)M.tri_off[m];
        const int T = M.ntris[m];
        Mat mt = load_mat(mat, i);
        V3 pmn = sub(p.pos, v3(r, r, r)), pmx = add(p.pos, v3(r, r, r));
        for (int w = 0; w < W.n_walls; ++w) {
            const int64_t idx = (int64_t)i * W.n_walls + w;
            // sphere test main.cpp:2318-2323
            V3 dp = sub(p.pos, ld3(W.pos + 3 * w));
            double d2 = dot(dp, dp);
            double rs = r + W.radius[w];
            int nc = 0;
            V3 ds = H.wprev_active[idx] ? ld3(H.wprev_ds + 3 * idx) : v3(0.0, 0.0, 0.0);
            if (!(d2 > rs * rs)) {
                const int t0 = W.tri_off[w], nt = W.ntri[w];
                for (int t = 0; t < nt; ++t) {
                    const double* wt9 = W.tris + 9 * (int64_t)(t0 + t);
                    if (!wall_tri_candidate(wt9, pmn, pmx)) continue;
                    V3 wt[3] = {ld3(wt9), ld3(wt9 + 3), ld3(wt9 + 6)};
                    double area;
                    V3 xc, cn;
                    int st = 0;
                    bool ok = wall_contact(wp, wv, ti, T, M.nverts[m], wt, ld3(W.normals + 3 * (int64_t)(t0 + t)), tol, PTS, PA,
                                           PB, pts_cap, area, xc, cn, st);
                    if (!ok) {
                        if (st && lane_id() == 0) atomicOr(status, (int)ST_WALL);
                        continue;
                    }
                    Contact c = pw_force(cn, area, xc, p, mt, W.mu[w], sp.dt, ds);
                    if (nc < W.max_wall_tris && lane_id() == 0) {
                        double* o = wcontrib + 6 * (idx * W.max_wall_tris + nc);
                        o[0] = c.f.x; o[1] = c.f.y; o[2] = c.f.z; o[3] = c.r1.x; o[4] = c.r1.y; o[5] = c.r1.z;
                    }
                    ++nc;
                }
            }
            if (lane_id() == 0) {
                wnc[idx] = nc;
                H.wcur_ds[3 * idx] = ds.x; H.wcur_ds[3 * idx + 1] = ds.y; H.wcur_ds[3 * idx + 2] = ds.z;
                H.wcur_active[idx] = nc > 0 ? 1 : 0;
            }
        }
        __syncthreads();
    }
}

__global__ void k_walls_clear(int64_t n, int32_t* wnc, uint8_t* wact, unsigned long long* stamp) {
    stamp_here(stamp);
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
        wnc[k] = 0;
        wact[k] = 0;
    }
}

// ------------------------------------------------------------ fold + integrate
// The step's counters summed from the per-pair and per-(particle, wall)
// arrays by nb workgroups of 256 threads (this one is bid): one atomic per
// value and workgroup into ctr, which k_report hands over and clears.
__device__ inline void sum_counters(const DPairs& P, const TopoBufs& T, int64_t n_pw, const int32_t* wnc,
                                    unsigned long long* ctr, int bid, int nb) {
    const int n = *P.n;
    unsigned long long c_loops = 0, c_checks = 0, c_seg = 0, c_comps = 0, c_raw = 0, c_tests = 0, c_hits = 0,
                       c_wall = 0, c_scan = 0, c_tp = 0, c_smax = 0;
    for (int p = bid * blockDim.x + threadIdx.x; p < n; p += nb * blockDim.x) {
        const unsigned long long ns = (unsigned long long)P.nseg[p];
        c_loops += T.ncontrib[p];
        c_checks += P.tri_checks[p];
        c_seg += ns;
        c_tp += ns > 0 ? 1ull : 0ull;
        c_smax = ns > c_smax ? ns : c_smax;
        c_comps += T.comps[p];
        c_raw += T.loops_raw[p];
        c_tests += P.tri_tests[p];
        c_hits += P.tri_hits[p];
        c_scan += P.tri_scanned[p];
    }
    for (int64_t k = bid * blockDim.x + threadIdx.x; k < n_pw; k += nb * blockDim.x) c_wall += wnc[k];
    unsigned long long v[11] = {c_loops, c_checks, c_seg, c_comps, c_raw, c_tests, c_hits, c_wall, c_scan, c_tp, c_smax};
    for (int k = 0; k < 10; ++k)
        for (int d = 32; d >= 1; d >>= 1) v[k] += __shfl_xor(v[k], d, WAVE);
    for (int d = 32; d >= 1; d >>= 1) {
        const unsigned long long o = __shfl_xor(v[10], d, WAVE);
        v[10] = o > v[10] ? o : v[10];
    }
    // combine the waves of the workgroup, then one atomic per value
    __shared__ unsigned long long s[11][16];
    __shared__ unsigned long long tot[11];
    const int w = threadIdx.x / WAVE, nw = blockDim.x / WAVE;
    if (lane_id() == 0)
        for (int k = 0; k < 11; ++k) s[k][w] = v[k];
    __syncthreads();
    if (threadIdx.x < 11) {
        const int k = threadIdx.x;
        unsigned long long r = s[k][0];
        for (int q = 1; q < nw; ++q) r = k == 10 ? (s[k][q] > r ? s[k][q] : r) : r + s[k][q];
        tot[k] = r;
    }
    __syncthreads();
    if (threadIdx.x < 11) {
        // v order: loops, checks, segments, comps, raw loops, tests, hits, wall contacts, scanned, pairs with
        // segments, largest segment count.  ctr: 0 contacts (accepted loops + wall contacts), 1 tri_checks,
        // 2 segments, 3 comps, 4 loops, 5 pairs, 6 wall contacts, 7 tri_tests, 8 tri_hits, 9 tri_scanned,
        // 10 pairs with segments, 11 largest segment count
        const int k = threadIdx.x;
        const int dst[11] = {0, 1, 2, 3, 4, 7, 8, 6, 9, 10, 11};
        if (k == 0) atomicAdd(&ctr[0], tot[0] + tot[7]);
        else if (k == 10) atomicMax(&ctr[11], tot[10]);
        else atomicAdd(&ctr[dst[k]], tot[k]);
    }
    if (bid == 0 && threadIdx.x == 0) atomicAdd(&ctr[5], (unsigned long long)n);
}

constexpr int FOLD_CTR_GROUPS = 128;
__global__ void __launch_bounds__(256) k_fold_integrate(int32_t n, const double* sin, double* sout, DMat mat,
                                                        const int32_t* adj_off, const int32_t* adj_j,
                                                        const int32_t* adj_pair, int32_t adj_cap, TopoBufs T,
                                                        int32_t n_walls,
                                                        const double* wcontrib, const int32_t* wnc, int32_t mwt,
                                                        double* Fo, double* To, int32_t* counts, SimParams sp,
                                                        DPairs P, unsigned long long* ctr,
                                                        unsigned long long* stamp_a, unsigned long long* stamp_b) {
    stamp_here(stamp_a, stamp_b);
    // the first FOLD_CTR_GROUPS workgroups also sum the step's counters (the
    // topology and the walls, which they read, are complete)
    if (ctr && blockIdx.x < FOLD_CTR_GROUPS)
        sum_counters(P, T, (int64_t)n * n_walls, wnc, ctr, blockIdx.x, min((int)gridDim.x, FOLD_CTR_GROUPS));
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    // the body and its inertia first: their loads overlap the fold's chains
    Body b = load_body(sin + 16 * (int64_t)i);
    M3 Ib;
    for (int r = 0; r < 3; ++r) for (int c = 0; c < 3; ++c) Ib.m[r][c] = mat.Ibinv[9 * (int64_t)i + 3 * r + c];
    const double im = mat.inv_mass[i];
    // forces[i] and torques[i] start at zero and receive, in the reference's
    // order (main.cpp:2049-2050 pair loop, then walls 2310-2447): pairs with
    // partners in ascending index, loops in order, then walls and their tris
    V3 F = v3(0.0, 0.0, 0.0), Tq = v3(0.0, 0.0, 0.0);
    int cnt = 0;
    if (adj_off && adj_off[n] <= adj_cap) {
        const int a0 = adj_off[i], a1 = adj_off[i + 1];
        for (int k = a0; k < a1; ++k) {
            const int j = adj_j[k];
            const int p = adj_pair[k];
            if (p < 0) continue;
            const int nl = T.ncontrib[p];
            const double* c = T.contrib + 9 * (int64_t)p * T.lmax;
            for (int l = 0; l < nl; ++l, c += 9) {
                V3 f = ld3(c);
                if (i < j) {
                    F = sub(F, f);
                    Tq = add(Tq, cross(ld3(c + 3), mul(f, -1.0)));
                } else {
                    F = add(F, f);
                    Tq = add(Tq, cross(ld3(c + 6), f));
                }
                ++cnt;
            }
        }
    }
    for (int w = 0; w < n_walls; ++w) {
        const int64_t idx = (int64_t)i * n_walls + w;
        const int nc = wnc[idx] < mwt ? wnc[idx] : mwt;
        for (int k = 0; k < nc; ++k) {
            const double* c = wcontrib + 6 * (idx * mwt + k);
            V3 f = ld3(c);
            F = add(F, f);
            Tq = add(Tq, cross(ld3(c + 3), f));
            ++cnt;
        }
    }
    Fo[3 * (int64_t)i] = F.x; Fo[3 * (int64_t)i + 1] = F.y; Fo[3 * (int64_t)i + 2] = F.z;
    To[3 * (int64_t)i] = Tq.x; To[3 * (int64_t)i + 1] = Tq.y; To[3 * (int64_t)i + 2] = Tq.z;
    counts[i] = cnt;
    integrate(b, Ib, im, F, Tq, v3(sp.g[0], sp.g[1], sp.g[2]), sp.dt);
    store_body(sout + 16 * (int64_t)i, b);
}

__global__ void k_iota_pairs(int32_t* pi, int32_t* pj, int32_t n) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) { pi[p] = 2 * p; pj[p] = 2 * p + 1; }
}


// Offsets of the per-pair call's packed output: one workgroup scans the
// pairs 1024 at a time (loops, loop vertices, segments of each pair) with a
// carry between chunks.
__global__ void __launch_bounds__(1024) k_api_scan(ApiPack k) {
    __shared__ long long sh[3][1024];
    __shared__ long long carry[3];
    const int t = threadIdx.x;
    if (t < 3) carry[t] = 0;
    __syncthreads();
    for (int base = 0; base < k.n; base += 1024) {
        const int p = base + t;
        long long v[3] = {0, 0, 0};
        if (p < k.n) {
            v[0] = k.nloops[p];
            for (int l = 0; l < k.nloops[p]; ++l) v[1] += (long long)k.rec[((int64_t)p * k.lmax + l) * 28 + 24];
            v[2] = k.nseg[p];
        }
        for (int c = 0; c < 3; ++c) sh[c][t] = v[c];
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            long long a[3];
            for (int c = 0; c < 3; ++c) a[c] = t >= off ? sh[c][t - off] : 0;
            __syncthreads();
            for (int c = 0; c < 3; ++c) sh[c][t] += a[c];
            __syncthreads();
        }
        if (p < k.n) {
            k.loop_off[p] = (int32_t)(carry[0] + sh[0][t] - v[0]);
            k.pvtx_off[p] = carry[1] + sh[1][t] - v[1];
            k.seg_out_off[p] = (int32_t)(carry[2] + sh[2][t] - v[2]);
        }
        __syncthreads();
        if (t < 3) carry[t] += sh[t][1023];
        __syncthreads();
    }
    if (t == 0) {
        k.loop_off[k.n] = (int32_t)carry[0];
        k.pvtx_off[k.n] = carry[1];
        k.seg_out_off[k.n] = (int32_t)carry[2];
        k.totals[0] = carry[0];
        k.totals[1] = carry[1];
        k.totals[2] = carry[2];
        if (k.vtx_off_out && carry[0] <= k.loop_cap) k.vtx_off_out[carry[0]] = (int32_t)carry[1];
    }
}

// One wave per pair moves its rows to the packed positions; a pair whose
// rows would pass an output's capacity writes nothing to that output (the
// host then reports the totals and the call fails with ZDX_CAPACITY).
__global__ void __launch_bounds__(256) k_api_pack(ApiPack k) {
    const int lane = threadIdx.x & 63;
    const int64_t nw = (int64_t)gridDim.x * (blockDim.x >> 6);
    for (int64_t p = blockIdx.x * (int64_t)(blockDim.x >> 6) + (threadIdx.x >> 6); p < k.n; p += nw) {
        const int nl = k.nloops[p];
        const int64_t lo = k.loop_off[p];
        if (lo + nl <= k.loop_cap) {
            if (k.rec_out)
                for (int64_t e = lane; e < 28 * (int64_t)nl; e += 64)
                    k.rec_out[28 * lo + e] = k.rec[(p * k.lmax) * 28 + e];
            if (k.ljac_out)
                for (int64_t e = lane; e < 72 * (int64_t)nl; e += 64)
                    k.ljac_out[72 * lo + e] = k.ljac[(p * k.lmax) * 72 + e];
        }
        if (k.vtx_out) {
            const int64_t vo = k.pvtx_off[p];
            const int64_t nv = k.pvtx_off[p + 1] - vo;
            if (lane == 0 && lo + nl <= k.loop_cap) {
                int64_t acc = vo;
                for (int l = 0; l < nl; ++l) {
                    k.vtx_off_out[lo + l] = (int32_t)acc;
                    acc += (int64_t)k.rec[(p * k.lmax + l) * 28 + 24];
                }
            }
            if (vo + nv <= k.vtx_cap)
                for (int64_t e = lane; e < 3 * nv; e += 64) k.vtx_out[3 * vo + e] = k.vtx[(p * k.vmax) * 3 + e];
        }
        if (k.seg_out) {
            const int64_t so = k.seg_out_off[p];
            const int ns = k.nseg[p];
            if (so + ns <= k.seg_cap)
                for (int64_t e = lane; e < 6 * (int64_t)ns; e += 64) k.seg_out[6 * so + e] = k.arena[6 * k.seg_off[p] + e];
        }
    }
}

__global__ void __launch_bounds__(1024) k_api_bodies(const zdx_pair_in* in, int32_t n, DMeshes M, int32_t* mesh,
                                                    double* state, int64_t* wvo, int64_t* wto) {
    __shared__ long long sh[2][1024];
    __shared__ long long carry[2];
    const int t = threadIdx.x;
    const int64_t nb = 2 * (int64_t)n;
    if (t < 2) carry[t] = 0;
    __syncthreads();
    for (int64_t base = 0; base < nb; base += 1024) {
        const int64_t b = base + t;
        long long v[2] = {0, 0};
        if (b < nb) {
            const zdx_pair_in& q = in[b >> 1];
            const int side = (int)(b & 1);
            const int m = side ? q.mesh_b : q.mesh_a;
            const double* pos = side ? q.pos_b : q.pos_a;
            const double* qu = side ? q.quat_b : q.quat_a;
            mesh[b] = m;
            double* st = state + 16 * b;
            st[0] = pos[0]; st[1] = pos[1]; st[2] = pos[2];
            st[3] = qu[0]; st[4] = qu[1]; st[5] = qu[2]; st[6] = qu[3];
            for (int k = 7; k < 16; ++k) st[k] = 0.0;
            v[0] = M.nverts[m];
            v[1] = M.ntris[m];
        }
        for (int c = 0; c < 2; ++c) sh[c][t] = v[c];
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            long long a[2];
            for (int c = 0; c < 2; ++c) a[c] = t >= off ? sh[c][t - off] : 0;
            __syncthreads();
            for (int c = 0; c < 2; ++c) sh[c][t] += a[c];
            __syncthreads();
        }
        if (b < nb) {
            wvo[b] = carry[0] + sh[0][t] - v[0];
            wto[b] = carry[1] + sh[1][t] - v[1];
        }
        __syncthreads();
        if (t < 2) carry[t] += sh[t][1023];
        __syncthreads();
    }
    if (t == 0) {
        wvo[nb] = carry[0];
        wto[nb] = carry[1];
    }
}

__global__ void __launch_bounds__(64) k_wall_query(DBodies B, DMeshes M, int32_t n_bodies, const double* tris,
                                                   const double* normals, int32_t n_tris, int32_t jac, int32_t cap,
                                                   double* scr, int32_t* isrc, double* out, int32_t* status) {
    LaneDev wp;
    const int64_t q = blockIdx.x;
    if (q >= (int64_t)n_bodies * n_tris) return;
    const int b = (int)(q / n_tris), t = (int)(q - (int64_t)b * n_tris);
    const int64_t per = 3 * (int64_t)cap + 4 * (int64_t)(cap + 8) + 28 * (int64_t)(cap + 8);
    SArr<double> PTS; PTS.p = scr + q * per; PTS.s = 1;
    SArr<double> PA = PTS.at(3 * (int64_t)cap);
    SArr<double> PB = PA.at(2 * (int64_t)(cap + 8));
    SArr<double> TA = PB.at(2 * (int64_t)(cap + 8));
    SArr<double> TB = TA.at(14 * (int64_t)(cap + 8));
    int32_t* src = isrc + q * (int64_t)(cap + 1);
    const int m = B.mesh[b];
    const double* wv = B.wv + 3 * B.wv_off[b];
    const int32_t* ti = M.tri_idx + 3 * (int64_t)M.tri_off[m];
    const double* wt9 = tris + 9 * (int64_t)t;
    V3 wt[3] = {ld3(wt9), ld3(wt9 + 3), ld3(wt9 + 6)};
    const V3 wn = ld3(normals + 3 * (int64_t)t);
    double area = 0.0;
    V3 xc, cn;
    int st = 0;
    const bool ok = wall_contact(wp, wv, ti, M.ntris[m], M.nverts[m], wt, wn, M.tol[m], PTS, PA, PB, cap, area, xc,
                                 cn, st, src);
    double* o = out + 32 * q;
    if (lane_id() != 0) return;
    for (int k = 0; k < 32; ++k) o[k] = 0.0;
    if (!ok) {
        if (st) atomicOr(status, 1);
        return;
    }
    const Body p = load_body(B.state + 16 * (int64_t)b);
    if (dot(cn, sub(p.pos, xc)) < 0.0) cn = mul(cn, -1.0);   // main.cpp:2363-2367
    o[0] = 1.0; o[1] = area;
    o[2] = xc.x; o[3] = xc.y; o[4] = xc.z;
    o[5] = cn.x; o[6] = cn.y; o[7] = cn.z;
    if (jac) {
        if (!wall_contact_jac(wv, ti, p.pos, wt, wn, M.tol[m], src, cap, TA, TB, o + 8, o + 14)) {
            atomicOr(status, 2);   // cannot differ from wall_contact: same doubles, same branches
        }
    }
}

__global__ void k_or_i32(const int32_t* v, int32_t n, int32_t* out) {
    int32_t acc = 0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        acc |= v[i];
    if (acc) atomicOr(out, acc);
}

inline int nblk(int64_t n, int bs) { return (int)((n + bs - 1) / bs); }

}  // namespace

#ifdef ZDX_WALK_PROFILE
extern "C" int zdx_topo_pairs(long long* out, int n) {
    return hipMemcpyFromSymbol(out, HIP_SYMBOL(g_topo_pairs), 4 * (size_t)n * sizeof(long long)) == hipSuccess ? 0 : 1;
}
extern "C" int zdx_narrow_pairs(long long* out, int n) {
    return hipMemcpyFromSymbol(out, HIP_SYMBOL(g_np_pairs), NP_REC * (size_t)n * sizeof(long long)) == hipSuccess ? 0 : 1;
}
extern "C" int zdx_walk_pairs(long long* out, int n) {
    return hipMemcpyFromSymbol(out, HIP_SYMBOL(g_walk_pairs), 12 * (size_t)n * sizeof(long long)) == hipSuccess ? 0 : 1;
}
extern "C" int zdx_walk_abs(long long* out, int n) {
    return hipMemcpyFromSymbol(out, HIP_SYMBOL(g_walk_abs), 32 * (size_t)n * sizeof(long long)) == hipSuccess ? 0 : 1;
}
extern "C" int zdx_snap_hits(long long* out, int n) {
    return hipMemcpyFromSymbol(out, HIP_SYMBOL(g_snap_hits), (size_t)n * sizeof(long long)) == hipSuccess ? 0 : 1;
}
extern "C" int zdx_walkd_pairs(long long* out, int n) {
    return hipMemcpyFromSymbol(out, HIP_SYMBOL(g_walkd_pairs), 8 * (size_t)n * sizeof(long long)) == hipSuccess ? 0 : 1;
}
extern "C" int zdx_walk_profile(unsigned long long* out) {
    if (hipMemcpyFromSymbol(out, HIP_SYMBOL(g_walk_prof), 16 * sizeof(unsigned long long)) != hipSuccess) return 1;
    unsigned long long z[16] = {0};
    return hipMemcpyToSymbol(HIP_SYMBOL(g_walk_prof), z, sizeof(z)) == hipSuccess ? 0 : 1;
}
#endif



// ------------------------------------------------------------ launchers
void launch_stamp(unsigned long long* slot, hipStream_t s) {
    if (slot) hipLaunchKernelGGL(k_stamp, dim3(1), dim3(1), 0, s, slot);
}
void launch_report(const int32_t* status, unsigned long long* ctr, const unsigned long long* kt, StepReport* rep,
                   hipStream_t s) {
    hipLaunchKernelGGL(k_report, dim3(1), dim3(64), 0, s, status, ctr, kt, rep);
}

void launch_transform(const DBodies& B, const DMeshes& M, hipStream_t s, const int32_t* list,
                      const int32_t* n_list, int32_t groups, unsigned long long* stamp) {
    if (B.n <= 0) return;
    int g = list ? groups : (B.n + 3) / 4;
    if (g < 1) g = 1;
    hipLaunchKernelGGL(k_transform, dim3(g), dim3(256), 0, s, B, M, list, n_list, stamp);
}

void launch_need_list(int32_t n, const double* state, const double* radius, const int32_t* cnt_all, const WallBufs& W,
                      int32_t* list, int32_t* n_list, int32_t* wlist, int32_t* n_wlist, bool zeroed, hipStream_t s,
                      unsigned long long* stamp) {
    if (!zeroed) {
        hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, n_list, 0, nullptr, (int64_t)1);
        hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, n_wlist, 0, nullptr, (int64_t)1);
    }
    hipLaunchKernelGGL(k_need_lists, dim3(nblk(n, 256)), dim3(256), 0, s, n, state, radius, cnt_all, W, list, n_list,
                       wlist, n_wlist, stamp);
}

void launch_broadphase(const BroadPhaseBufs& bp, hipStream_t s, const NeedLists* need, unsigned long long* kt0,
                       const StepZero* z) {
    NeedLists nl{};
    if (need) nl = *need;
    StepZero sz{};
    if (kt0 && z) sz = *z;
    const int n = bp.n;
    hipLaunchKernelGGL(k_bp_bounds_part, dim3(BP_PARTS), dim3(256), 0, s, bp.state, bp.radius, n, bp.part,
                       z ? kt0 : nullptr, sz);
    // k_bp_cells writes the grid and its cell count (to a scalar, n_adj,
    // written again at the end) for the cell scan
    const int32_t* ncells = bp.n_adj;
    hipLaunchKernelGGL(k_bp_cells, dim3(std::max(1, nblk(n, 256))), dim3(256), 0, s, bp);
    scan_i32(bp.cell_count, ncells, bp.max_cells, bp.scan_tmp, bp.cell_start, nullptr, s);
    hipLaunchKernelGGL(k_bp_scatter, dim3(nblk(n, 256)), dim3(256), 0, s, bp);
    hipLaunchKernelGGL(k_bp_count, dim3(nblk(n, 256)), dim3(256), 0, s, bp, nl);
    scan_i32x2(bp.cnt_up, bp.cnt_all, n, bp.scan_tmp, bp.pair_off, bp.adj_off, s);
    hipLaunchKernelGGL(k_bp_fill, dim3(nblk(n, 128)), dim3(128), 0, s, bp);
    hipLaunchKernelGGL(k_bp_adjfix, dim3(nblk(n, 256) > 0 ? nblk(n, 256) : 1), dim3(256), 0, s, bp);
}

// the workgroup kernel over the heavy list left by the wave kernel
template <int BS, int K>
static void narrow_t(const DBodies& B, const DMeshes& M, const DPairs& P, const NarrowCfg& cfg, int32_t* li, float* lf,
                     int32_t tmax, unsigned int* q, hipStream_t s) {
    hipLaunchKernelGGL((k_narrow<BS, K>), dim3(cfg.groups), dim3(BS), (size_t)narrow_lds_bytes(BS), s, B, M, P, li,
                       lf, tmax, cfg.use_queue, q + NQ_WG, P.heavy, q + NQ_HEAVY, q + NQ_HEAVY_BACK, nullptr);
}

void launch_narrow(const DBodies& B, const DMeshes& M, const DPairs& P, const NarrowCfg& cfg, int32_t* li, float* lf,
                   int32_t seg_cap, int32_t tmax, unsigned int* q, hipStream_t s, bool zeroed,
                   unsigned long long* stamp) {
    if (!zeroed) {
        hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, (int32_t*)q, 0, nullptr, (int64_t)NQ_COUNTERS);
        hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, s, P.hit_top, 0ull, (int64_t)HIT_SHARDS * HIT_TOP_STRIDE);
    }
    // every pair through the one-wave kernel; the pairs it leaves (heavy)
    // through the workgroup kernel of the configured shape
    const int wg = cfg.wave_groups > 0 ? cfg.wave_groups : std::max(1, cfg.groups);
    const int32_t wt = cfg.wave_tests > 0 ? cfg.wave_tests : (int32_t)W_HEAVY_TESTS;
    hipLaunchKernelGGL(k_narrow_wave<4>, dim3(wg), dim3(NWV * WAVE), (size_t)narrow_wave_lds_bytes(), s, B, M, P,
                       P.heavy, q, cfg.use_queue, wt, stamp);
    const int key = P.prov ? 25604 : cfg.block * 100 + cfg.items;
    switch (key) {
        case 6404: narrow_t<64, 4>(B, M, P, cfg, li, lf, tmax, q, s); break;
        case 6408: narrow_t<64, 8>(B, M, P, cfg, li, lf, tmax, q, s); break;
        case 12804: narrow_t<128, 4>(B, M, P, cfg, li, lf, tmax, q, s); break;
        case 12808: narrow_t<128, 8>(B, M, P, cfg, li, lf, tmax, q, s); break;
        case 25601: narrow_t<256, 1>(B, M, P, cfg, li, lf, tmax, q, s); break;
        case 25602: narrow_t<256, 2>(B, M, P, cfg, li, lf, tmax, q, s); break;
        case 25604: narrow_t<256, 4>(B, M, P, cfg, li, lf, tmax, q, s); break;
        case 25608: narrow_t<256, 8>(B, M, P, cfg, li, lf, tmax, q, s); break;
        case 51202: narrow_t<512, 2>(B, M, P, cfg, li, lf, tmax, q, s); break;
        case 51204: narrow_t<512, 4>(B, M, P, cfg, li, lf, tmax, q, s); break;
        default: narrow_t<256, 4>(B, M, P, cfg, li, lf, tmax, q, s); break;
    }
    const int ge = std::max(1, cfg.groups) * 4;   // grid-stride over the hits
    const int gc = std::max(1, cfg.groups);        // 4 waves per workgroup over the listed pairs
    if (P.prov) {
        hipLaunchKernelGGL(k_narrow_exact<true>, dim3(ge), dim3(256), 0, s, B, P);
        hipLaunchKernelGGL(k_narrow_compact<true>, dim3(gc), dim3(256), 0, s, P, seg_cap);
    } else {
        hipLaunchKernelGGL(k_narrow_exact<false>, dim3(ge), dim3(256), 0, s, B, P);
        hipLaunchKernelGGL(k_narrow_compact<false>, dim3(gc), dim3(256), 0, s, P, seg_cap);
    }
    hipLaunchKernelGGL(k_topo_lists, dim3(gc), dim3(256), 0, s, P);
}

void launch_lpt_order(const DPairs& P, const DMeshes& M, const DBodies& B, int32_t* order, int32_t* bcnt,
                      int32_t* bcur, int32_t pair_cap, hipStream_t s, bool zeroed, unsigned long long* stamp,
                      const TopoBufs* tz, uint8_t* cur_active) {
    TopoBufs T{};
    if (tz) T = *tz;
    if (!zeroed) {
        hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, bcnt, 0, nullptr, (int64_t)64);
        hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, bcur, 0, nullptr, (int64_t)64);
    }
    const int g = std::max(1, std::min(nblk(pair_cap, 256), 512));
    hipLaunchKernelGGL(k_lpt_count, dim3(g), dim3(256), 0, s, P, M, B, bcnt, stamp);
    hipLaunchKernelGGL(k_lpt_scatter, dim3(g), dim3(256), 0, s, P, M, B, (const int32_t*)bcnt, bcur, order, T,
                       tz ? cur_active : nullptr, tz ? 1 : 0);
}

void launch_topo_zero(const DPairs& P, const TopoBufs& T, uint8_t* cur_active, int32_t pair_cap, hipStream_t s,
                      unsigned long long* stamp) {
    int g = nblk(pair_cap, 256);
    if (g > 2048) g = 2048;
    if (g < 1) g = 1;
    hipLaunchKernelGGL(k_topo_zero, dim3(g), dim3(256), 0, s, P, T, cur_active, stamp);
}

static void launch_pass_a(const DPairs& P, const DMeshes& M, const DBodies& B, const TopoBufs& T, hipStream_t s,
                          bool zeroed = false, hipStream_t side = nullptr, hipEvent_t ev_fork = nullptr,
                          hipEvent_t ev_join = nullptr) {
    if (!zeroed) launch_fill_u64(T.top, 0ull, 5, s);
    hipLaunchKernelGGL(k_tp_alloc, dim3((unsigned)((T.pass_groups + 255) / 256 * 4)), dim3(256), 0, s, P, T);
    if (P.topo_big) {
        // the pairs with few segments in small workgroups (many per CU) and
        // the big ones with their extra units in 1024-thread workgroups; with
        // a side stream the two launches run side by side
        const bool fork = side && ev_fork && ev_join;
        hipStream_t sb = fork ? side : s;
        if (fork) {
            (void)hipEventRecord(ev_fork, s);
            (void)hipStreamWaitEvent(side, ev_fork, 0);
        }
        hipLaunchKernelGGL((k_tp_pass_a<TP_PASS_A_THREADS, TP_PASS_A_KEYS, TP_PASS_A_WORDS, 2>),
                           dim3((unsigned)T.pass_groups), dim3(TP_PASS_A_THREADS),
                           84 * TP_PASS_A_SEGS + TP_PASS_A_KEYS + 8 * TP_PASS_A_WORDS, sb, P, M, B, T,
                           (int32_t)TP_PASS_A_SEGS);
        hipLaunchKernelGGL(k_tp_snap_cc, dim3(256), dim3(CC_THREADS), 0, sb, P, T);
        if (fork) (void)hipEventRecord(ev_join, side);
        constexpr int SK = TOPO_BIG_SEGS * TOPO_BIG_SEGS, SW = 64;
        hipLaunchKernelGGL((k_tp_pass_a<256, SK, SW, 1>), dim3((unsigned)T.pass_groups * 4), dim3(256),
                           84 * TOPO_BIG_SEGS + SK + 8 * SW, s, P, M, B, T, (int32_t)TOPO_BIG_SEGS);
        if (fork) (void)hipStreamWaitEvent(s, ev_join, 0);
    } else {
        hipLaunchKernelGGL((k_tp_pass_a<TP_PASS_A_THREADS, TP_PASS_A_KEYS, TP_PASS_A_WORDS, 0>),
                           dim3((unsigned)T.pass_groups), dim3(TP_PASS_A_THREADS),
                           84 * TP_PASS_A_SEGS + TP_PASS_A_KEYS + 8 * TP_PASS_A_WORDS, s, P, M, B, T,
                           (int32_t)TP_PASS_A_SEGS);
    }
}

void launch_topo_sim(const DPairs& P, const DMeshes& M, const DBodies& B, const DMat& mat, const HistBufs& H,
                     const TopoBufs& T, const SimParams& sp, hipStream_t s, bool zeroed, hipStream_t side,
                     hipEvent_t ev_fork, hipEvent_t ev_join, hipEvent_t ev_cc) {
    constexpr int WALK_BIG_GROUPS = 128;   // the big list is short (pairs of 40 or more segments)
    if (!(side && ev_fork && ev_join && ev_cc && P.topo_big) || T.nthreads < 2 * WALK_BIG_GROUPS) {
        launch_pass_a(P, M, B, T, s, zeroed, side, ev_fork, ev_join);
        hipLaunchKernelGGL(k_tp_walk_bcd_sim<1>, dim3((unsigned)T.nthreads), dim3(64), T.lds_bytes, s, P, M, B, mat,
                           H, T, sp, T.lds_bytes, 0);
        return;
    }
    // the big pairs (the long chains) on the side stream: pass A, their snap
    // clusters and their walks, which so start without waiting for pass A of
    // the small pairs; those and their walks on the main stream.  The small
    // pairs' walks also wait for the snap clusters: a walk wave (over 256
    // VGPRs) leaves too few registers on its SIMD for the multi-wave
    // workgroups of pass A and k_tp_snap_cc, which would otherwise wait for
    // the walks to end.  The two walk launches use disjoint slices of the
    // global walk scratch.
    if (!zeroed) launch_fill_u64(T.top, 0ull, 5, s);
    hipLaunchKernelGGL(k_tp_alloc, dim3((unsigned)((T.pass_groups + 255) / 256 * 4)), dim3(256), 0, s, P, T);
    (void)hipEventRecord(ev_fork, s);
    (void)hipStreamWaitEvent(side, ev_fork, 0);
    hipLaunchKernelGGL((k_tp_pass_a<TP_PASS_A_THREADS, TP_PASS_A_KEYS, TP_PASS_A_WORDS, 2>),
                       dim3((unsigned)T.pass_groups), dim3(TP_PASS_A_THREADS),
                       84 * TP_PASS_A_SEGS + TP_PASS_A_KEYS + 8 * TP_PASS_A_WORDS, side, P, M, B, T,
                       (int32_t)TP_PASS_A_SEGS);
    hipLaunchKernelGGL(k_tp_snap_cc, dim3(256), dim3(CC_THREADS), 0, side, P, T);
    (void)hipEventRecord(ev_cc, side);
    const int64_t gs = T.nthreads - WALK_BIG_GROUPS;
    TopoBufs Tb = T;
    Tb.scr_i = T.scr_i + gs * T.scr_ni;
    Tb.scr_d = T.scr_d + gs * T.scr_nd;
    hipLaunchKernelGGL(k_tp_walk_bcd_sim<1>, dim3((unsigned)WALK_BIG_GROUPS), dim3(64), T.lds_bytes, side, P, M, B,
                       mat, H, Tb, sp, T.lds_bytes, 2);
    (void)hipEventRecord(ev_join, side);
    constexpr int SK = TOPO_BIG_SEGS * TOPO_BIG_SEGS, SW = 64;
    hipLaunchKernelGGL((k_tp_pass_a<256, SK, SW, 1>), dim3((unsigned)T.pass_groups * 4), dim3(256),
                       84 * TOPO_BIG_SEGS + SK + 8 * SW, s, P, M, B, T, (int32_t)TOPO_BIG_SEGS);
    // the small pairs' walks with the LDS their largest pair needs, so that
    // many more of them fit a CU at once
    const int Ss = TOPO_BIG_SEGS - 1;
    int64_t need = std::max(tp_scratch_bytes(tp_walk_b_ints(Ss, T.adj_row), tp_walk_b_dbls(Ss)),
                            tp_scratch_bytes(tp_walk_d_ints(Ss), tp_walk_d_dbls(Ss)));
    need = std::min<int64_t>((need + 1023) & ~(int64_t)1023, T.lds_bytes);
    (void)hipStreamWaitEvent(s, ev_cc, 0);
    hipLaunchKernelGGL(k_tp_walk_bcd_sim<2>, dim3((unsigned)gs), dim3(64), (size_t)need, s, P, M, B, mat, H, T, sp,
                       (int32_t)need, 1);
    (void)hipStreamWaitEvent(s, ev_join, 0);
}
void launch_topo_api(const DPairs& P, const DMeshes& M, const DBodies& B, const TopoBufs& T, int32_t split,
                     hipStream_t s) {
    if (T.ej) hipLaunchKernelGGL(k_tp_endpoint_jac, dim3((unsigned)T.pass_groups), dim3(256), 0, s, P, M, B, T);
    launch_pass_a(P, M, B, T, s);
    hipLaunchKernelGGL(k_tp_walk_b, dim3((unsigned)T.nthreads), dim3(64), T.lds_bytes, s, P, M, B, T, split,
                       T.lds_bytes);
    hipLaunchKernelGGL(k_tp_pass_c, dim3((unsigned)T.pass_groups), dim3(256), 0, s, P, M, B, T);
    hipLaunchKernelGGL(k_tp_walk_d_api, dim3((unsigned)T.nthreads), dim3(64), T.lds_bytes, s, P, M, B, T,
                       T.lds_bytes);
}
void launch_walls(const DBodies& B, const DMeshes& M, const DMat& mat, const WallBufs& W, const HistBufs& H,
                  double* wcontrib, int32_t* wnc, const int32_t* list, const int32_t* n_list, double* scr,
                  int64_t groups, int32_t pts_cap, const SimParams& sp, int32_t* status, hipStream_t s,
                  unsigned long long* stamp) {
    if (W.n_walls <= 0) return;
    const int64_t n = (int64_t)B.n * W.n_walls;
    hipLaunchKernelGGL(k_walls_clear, dim3(nblk(n, 256) < 4096 ? nblk(n, 256) : 4096), dim3(256), 0, s, n, wnc,
                       H.wcur_active, stamp);
    hipLaunchKernelGGL(k_walls, dim3((unsigned)groups), dim3(64), 0, s, B, M, mat, W, H, wcontrib, wnc, list, n_list,
                       scr, pts_cap, sp, status);
}
void launch_fold_integrate(int32_t n, const double* sin, double* sout, const DMat& mat, const int32_t* adj_off,
                           const int32_t* adj_j, const int32_t* adj_pair, int32_t adj_cap, const TopoBufs& T,
                           int32_t n_walls,
                           const double* wcontrib, const int32_t* wnc, int32_t mwt, double* F, double* Tq,
                           int32_t* counts, const SimParams& sp, const DPairs& P, unsigned long long* ctr,
                           hipStream_t s, unsigned long long* stamp_a, unsigned long long* stamp_b) {
    hipLaunchKernelGGL(k_fold_integrate, dim3(nblk(n, 256)), dim3(256), 0, s, n, sin, sout, mat, adj_off, adj_j,
                       adj_pair, adj_cap, T, n_walls, wcontrib, wnc, mwt, F, Tq, counts, sp, P, ctr, stamp_a,
                       stamp_b);
}
void launch_fill_i32(int32_t* p, int32_t v, int64_t n, hipStream_t s) {
    if (n <= 0) return;
    int g = nblk(n, 256);
    if (g > 4096) g = 4096;
    hipLaunchKernelGGL(k_fill_i32, dim3(g), dim3(256), 0, s, p, v, nullptr, n);
}
void launch_fill_u64(unsigned long long* p, unsigned long long v, int64_t n, hipStream_t s) {
    if (n <= 0) return;
    int g = nblk(n, 256);
    if (g > 4096) g = 4096;
    hipLaunchKernelGGL(k_fill_u64, dim3(g), dim3(256), 0, s, p, v, n);
}
void launch_iota_pairs(int32_t* pi, int32_t* pj, int32_t n, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_iota_pairs, dim3(nblk(n, 256)), dim3(256), 0, s, pi, pj, n);
}

void launch_api_pack(const ApiPack& k, hipStream_t s) {
    if (k.n <= 0) return;
    hipLaunchKernelGGL(k_api_scan, dim3(1), dim3(1024), 0, s, k);
    hipLaunchKernelGGL(k_api_pack, dim3(std::min(nblk(k.n, 4), 4096)), dim3(256), 0, s, k);
}

void launch_api_bodies(const void* in, int32_t n, const DMeshes& M, int32_t* mesh, double* state, int64_t* wv_off,
                       int64_t* wt_off, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_api_bodies, dim3(1), dim3(1024), 0, s, (const zdx_pair_in*)in, n, M, mesh, state, wv_off,
                       wt_off);
}

void launch_or_i32(const int32_t* v, int32_t n, int32_t* out, hipStream_t s) {
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, out, 0, nullptr, (int64_t)1);
    if (n > 0) hipLaunchKernelGGL(k_or_i32, dim3(std::min(nblk(n, 256), 1024)), dim3(256), 0, s, v, n, out);
}

void launch_wall_query(const DBodies& B, const DMeshes& M, int32_t n_bodies, const double* tris, const double* normals,
                       int32_t n_tris, int32_t jac, int32_t cap, double* scr, int32_t* isrc, double* out,
                       int32_t* status, hipStream_t s) {
    const int64_t nq = (int64_t)n_bodies * n_tris;
    if (nq <= 0) return;
    hipLaunchKernelGGL(k_wall_query, dim3((unsigned)nq), dim3(64), 0, s, B, M, n_bodies, tris, normals, n_tris, jac,
                       cap, scr, isrc, out, status);
}

}  // namespace zdx
