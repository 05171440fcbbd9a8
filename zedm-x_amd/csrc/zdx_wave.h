// zdx_wave.h — lane, wave and workgroup primitives of the HIP kernels.
//
// Textually included inside namespace zdx::{anonymous} after WAVE (64) is
// defined: by zdx_kernels.hip for gfx950, and by the host SIMT emulation of
// the narrow kernel (tests/csrc/simt_emu.h), which supplies the same
// intrinsics (ballot, lane shuffles, readfirstlane, barriers) for CPU runs.
// No includes or namespaces of its own.

// the workgroup's dynamic LDS, as a byte array
#ifndef ZDX_DYN_LDS
#define ZDX_DYN_LDS(name) extern __shared__ char name[]
#endif

__device__ inline int lane_id() { return __lane_id(); }

// Exclusive rank of `pred` among the workgroup's threads (thread order) and
// the total count.  Uses wave ballots and one LDS array of per-wave counts.
template <int BS>
__device__ inline int block_rank(bool pred, int* s_wave, int& total) {
    constexpr int NW = BS / WAVE;
    const int lane = lane_id();
    const int wid = threadIdx.x / WAVE;
    unsigned long long m = __ballot(pred);
    unsigned long long below = (lane == 0) ? 0ull : (m & ((~0ull) >> (64 - lane)));
    int r = __popcll(below);
    if (lane == 0) s_wave[wid] = __popcll(m);
    __syncthreads();
    int off = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
        int c = s_wave[w];
        off += (w < wid) ? c : 0;
        tot += c;
    }
    __syncthreads();
    total = tot;
    return off + r;
}

// Exclusive prefix of per-thread counts (thread order) and the total.
template <int BS>
__device__ inline int block_excl(int v, int* s_wave, int& total) {
    constexpr int NW = BS / WAVE;
    const int lane = lane_id();
    const int wid = threadIdx.x / WAVE;
    int x = v;
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1) {
        int y = __shfl_up(x, d, WAVE);
        if (lane >= d) x += y;
    }
    if (lane == WAVE - 1) s_wave[wid] = x;
    __syncthreads();
    int off = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
        int c = s_wave[w];
        off += (w < wid) ? c : 0;
        tot += c;
    }
    __syncthreads();
    total = tot;
    return off + x - v;
}

__device__ inline V3 ld3(const double* p) { return v3(p[0], p[1], p[2]); }
// A value that every lane of the wave holds, marked as such (scalar register).
__device__ inline int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ inline float uni(float v) { return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v))); }

// Stage stamp from inside the kernel that starts a stage (workgroup 0, thread
// 0, at its start): the device's constant-rate clock into one or two slots,
// which saves a stamp launch between the stages.
__device__ inline void stamp_here(unsigned long long* a, unsigned long long* b = nullptr) {
    if ((a || b) && blockIdx.x == 0 && threadIdx.x == 0) {
        const unsigned long long t = (unsigned long long)wall_clock64();
        if (a) *a = t;
        if (b) *b = t;
    }
}
