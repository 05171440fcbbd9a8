// simt_emu.h — host emulation of one HIP workgroup at a time, for CPU tests
// of device source (zedm-x_amd/csrc/zdx_wave.h, zdx_narrow.h).
//
// Every thread of the workgroup runs as a fiber (ucontext) on its own stack.
// The scheduler runs each runnable fiber until it blocks, then resolves
//   - a wave operation (ballot, lane shuffles, readfirstlane) once all 64
//     lanes of the wave wait at the same call site after the same number of
//     wave operations; any other state of the wave is reported as divergence
//     (on the GPU a lane shuffle that reads an inactive lane returns 0, so
//     device code must issue these with the whole wave active);
//   - a barrier once every thread of the workgroup waits at one.
// A wave barrier (__builtin_amdgcn_wave_barrier) is a wave operation with no
// data: it holds each lane until the whole wave has arrived, which is what
// lockstep execution gives the device code.
// readfirstlane is checked to see one value in all lanes (the device source
// uses it only to mark uniform values).  The workgroups of a grid run one
// after the other; __shared__ variables become statics, the dynamic LDS is a
// heap block of exactly the launch's size filled with a byte pattern (so
// AddressSanitizer sees an overflow and stale contents stay visible), and
// atomics are plain read-modify-writes.
#pragma once
#include <ucontext.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <memory>
#include <string>
#include <vector>

#if defined(__SANITIZE_ADDRESS__)
#include <sanitizer/common_interface_defs.h>
#define SIMT_ASAN 1
#endif

#define __global__
#define __device__
#define __host__
#define __shared__ static
#define __launch_bounds__(...)
#define ZDX_HD
#define ZDX_DYN_LDS(name) char* name = simt::dyn_lds()
#define __builtin_amdgcn_readfirstlane(v) simt::readfirstlane(v)
#define __builtin_amdgcn_wave_barrier() simt::wave_barrier()
#define __builtin_amdgcn_fence(order, scope) ((void)0)
#define threadIdx (simt::thread_idx())
#define blockIdx (simt::block_idx())
#define blockDim (simt::block_dim())
#define gridDim (simt::grid_dim())

struct int2 { int x, y; };
struct int4 { int x, y, z, w; };
inline int2 make_int2(int x, int y) { return int2{x, y}; }
inline int4 make_int4(int x, int y, int z, int w) { return int4{x, y, z, w}; }

namespace simt {

constexpr int W = 64;
constexpr size_t STACK = 256 * 1024;
enum State { RUN, BARRIER, WAVEOP, DONE };
enum Op { BALLOT, SHFL, SHFL_UP, SHFL_XOR, RFL, WSYNC };
struct Dim { unsigned x, y, z; };

struct Fiber {
    ucontext_t ctx;
    std::unique_ptr<char[]> stack;
    State st = RUN;
    Op op = BALLOT;
    const void* site = nullptr;
    long nops = 0;
    uint64_t val = 0, res = 0;
    int arg = 0;
    void* fake = nullptr;
};

struct Block {
    std::vector<Fiber> f;
    int cur = -1;
    unsigned bs = 0, bx = 0, gx = 0;
    std::unique_ptr<char[]> lds;
    ucontext_t main;
    void* main_fake = nullptr;
    const void* main_bottom = nullptr;
    size_t main_size = 0;
    std::function<void()> body;
    std::string error;
    long barriers = 0, waveops = 0;
};
inline Block& blk() {
    static Block b;
    return b;
}

inline Dim thread_idx() { return Dim{(unsigned)blk().cur, 0, 0}; }
inline Dim block_idx() { return Dim{blk().bx, 0, 0}; }
inline Dim block_dim() { return Dim{blk().bs, 1, 1}; }
inline Dim grid_dim() { return Dim{blk().gx, 1, 1}; }
inline char* dyn_lds() { return blk().lds.get(); }

inline void to_main() {
    Block& b = blk();
    Fiber& me = b.f[b.cur];
#ifdef SIMT_ASAN
    __sanitizer_start_switch_fiber(&me.fake, b.main_bottom, b.main_size);
#endif
    swapcontext(&me.ctx, &b.main);
#ifdef SIMT_ASAN
    __sanitizer_finish_switch_fiber(me.fake, nullptr, nullptr);
#endif
}

inline void fiber_entry() {
    Block& b = blk();
#ifdef SIMT_ASAN
    __sanitizer_finish_switch_fiber(nullptr, &b.main_bottom, &b.main_size);
#endif
    b.body();
    b.f[b.cur].st = DONE;
#ifdef SIMT_ASAN
    __sanitizer_start_switch_fiber(nullptr, b.main_bottom, b.main_size);
#endif
    // returning resumes uc_link (the scheduler)
}

inline void run_fiber(int t) {
    Block& b = blk();
    b.cur = t;
    Fiber& f = b.f[t];
#ifdef SIMT_ASAN
    __sanitizer_start_switch_fiber(&b.main_fake, f.stack.get(), STACK);
#endif
    swapcontext(&b.main, &f.ctx);
#ifdef SIMT_ASAN
    __sanitizer_finish_switch_fiber(b.main_fake, nullptr, nullptr);
#endif
}

inline uint64_t waveop(Op op, uint64_t v, int arg, const void* site) {
    Block& b = blk();
    Fiber& me = b.f[b.cur];
    me.st = WAVEOP;
    me.op = op;
    me.val = v;
    me.arg = arg;
    me.site = site;
    ++me.nops;
    to_main();
    return me.res;
}

inline const char* state_name(State s) {
    return s == RUN ? "run" : s == BARRIER ? "barrier" : s == WAVEOP ? "wave-op" : "done";
}

// resolves the wave operation of wave w when its lanes are all waiting at
// the same one; returns false when the wave is not in that state
inline bool resolve_wave(int w) {
    Block& b = blk();
    const int l0 = w * W;
    Fiber* L = &b.f[l0];
    for (int l = 0; l < W; ++l)
        if (L[l].st != WAVEOP) return false;
    for (int l = 1; l < W; ++l) {
        if (L[l].site != L[0].site || L[l].nops != L[0].nops || L[l].op != L[0].op) {
            char buf[256];
            std::snprintf(buf, sizeof buf, "wave %d: lanes 0 and %d wait at different wave operations "
                          "(op %d / %d, count %ld / %ld)", w, l, (int)L[0].op, (int)L[l].op, L[0].nops, L[l].nops);
            b.error = buf;
            return false;
        }
    }
    uint64_t r[W];
    switch (L[0].op) {
        case BALLOT: {
            uint64_t m = 0;
            for (int l = 0; l < W; ++l) m |= (uint64_t)(L[l].val & 1u) << l;
            for (int l = 0; l < W; ++l) r[l] = m;
            break;
        }
        case SHFL:
            for (int l = 0; l < W; ++l) r[l] = L[((L[l].arg % W) + W) % W].val;
            break;
        case SHFL_UP:
            for (int l = 0; l < W; ++l) r[l] = l >= L[l].arg ? L[l - L[l].arg].val : L[l].val;
            break;
        case SHFL_XOR:
            for (int l = 0; l < W; ++l) r[l] = L[(l ^ L[l].arg) & (W - 1)].val;
            break;
        case WSYNC:
            for (int l = 0; l < W; ++l) r[l] = 0;
            break;
        case RFL:
            for (int l = 0; l < W; ++l) {
                if (L[l].val != L[0].val) {
                    char buf[160];
                    std::snprintf(buf, sizeof buf, "wave %d: readfirstlane of a value that differs between lanes "
                                  "0 and %d", w, l);
                    b.error = buf;
                    return false;
                }
                r[l] = L[0].val;
            }
            break;
    }
    for (int l = 0; l < W; ++l) {
        L[l].res = r[l];
        L[l].st = RUN;
    }
    ++b.waveops;
    return true;
}

// Runs body() as a grid of `grid` workgroups of `block` threads (block a
// multiple of 64) with lds_bytes of dynamic LDS each.  Returns "" or the
// first error (divergence, barrier with exited threads, deadlock).
inline std::string launch(unsigned grid, unsigned block, size_t lds_bytes, std::function<void()> body) {
    Block& b = blk();
    if (block == 0 || block % W != 0) return "block size must be a multiple of 64";
    b.body = std::move(body);
    b.bs = block;
    b.gx = grid;
    b.error.clear();
    for (unsigned g = 0; g < grid; ++g) {
        b.bx = g;
        b.lds.reset(new char[lds_bytes > 0 ? lds_bytes : 1]);
        std::memset(b.lds.get(), 0xA5, lds_bytes > 0 ? lds_bytes : 1);
        b.f.clear();
        b.f.resize(block);
        for (unsigned t = 0; t < block; ++t) {
            Fiber& f = b.f[t];
            f.stack.reset(new char[STACK]);
            getcontext(&f.ctx);
            f.ctx.uc_stack.ss_sp = f.stack.get();
            f.ctx.uc_stack.ss_size = STACK;
            f.ctx.uc_link = &b.main;
            makecontext(&f.ctx, (void (*)())fiber_entry, 0);
        }
        for (;;) {
            for (unsigned t = 0; t < block; ++t)
                if (b.f[t].st == RUN) run_fiber((int)t);
            unsigned done = 0, bar = 0;
            for (unsigned t = 0; t < block; ++t) {
                done += b.f[t].st == DONE;
                bar += b.f[t].st == BARRIER;
            }
            if (done == block) break;
            bool progress = false;
            for (unsigned w = 0; w < block / W; ++w) {
                progress |= resolve_wave((int)w);
                if (!b.error.empty()) return b.error;
            }
            if (progress) continue;
            if (bar == block) {
                for (unsigned t = 0; t < block; ++t) b.f[t].st = RUN;
                ++b.barriers;
                continue;
            }
            // nothing can move: report the states of the first stuck wave
            std::string e = "workgroup " + std::to_string(g) + " cannot proceed:";
            for (unsigned w = 0; w < block / W; ++w) {
                int n[4] = {0, 0, 0, 0};
                for (int l = 0; l < W; ++l) ++n[b.f[w * W + l].st];
                e += " wave " + std::to_string(w) + " (barrier " + std::to_string(n[BARRIER]) + ", wave-op " +
                     std::to_string(n[WAVEOP]) + ", done " + std::to_string(n[DONE]) + ")";
            }
            return e;
        }
    }
    b.f.clear();
    return "";
}

template <class T>
inline uint64_t to_bits(T v) {
    static_assert(sizeof(T) <= 8, "64-bit values at most");
    uint64_t u = 0;
    std::memcpy(&u, &v, sizeof(T));
    return u;
}
template <class T>
inline T from_bits(uint64_t u) {
    T v;
    std::memcpy(&v, &u, sizeof(T));
    return v;
}

template <class T>
__attribute__((noinline)) T readfirstlane(T v) {
    return from_bits<T>(waveop(RFL, to_bits(v), 0, __builtin_return_address(0)));
}
// a wave-level barrier: all 64 lanes must arrive (device code issues it where
// lanes read LDS that other lanes of the wave wrote)
__attribute__((noinline)) inline void wave_barrier() { waveop(WSYNC, 0, 0, __builtin_return_address(0)); }

}  // namespace simt

// ---- the HIP intrinsics the device source uses
__attribute__((noinline)) inline unsigned long long __ballot(int pred) {
    return simt::waveop(simt::BALLOT, pred != 0, 0, __builtin_return_address(0));
}
template <class T>
__attribute__((noinline)) T __shfl(T v, int src, int width = 64) {
    if (width != 64) std::abort();
    return simt::from_bits<T>(simt::waveop(simt::SHFL, simt::to_bits(v), src, __builtin_return_address(0)));
}
template <class T>
__attribute__((noinline)) T __shfl_up(T v, unsigned d, int width = 64) {
    if (width != 64) std::abort();
    return simt::from_bits<T>(simt::waveop(simt::SHFL_UP, simt::to_bits(v), (int)d, __builtin_return_address(0)));
}
template <class T>
__attribute__((noinline)) T __shfl_xor(T v, int m, int width = 64) {
    if (width != 64) std::abort();
    return simt::from_bits<T>(simt::waveop(simt::SHFL_XOR, simt::to_bits(v), m, __builtin_return_address(0)));
}
inline void __syncthreads() {
    simt::Block& b = simt::blk();
    b.f[b.cur].st = simt::BARRIER;
    simt::to_main();
}
inline int __lane_id() { return simt::blk().cur % simt::W; }
inline int __popcll(unsigned long long x) { return __builtin_popcountll(x); }
inline unsigned atomicAdd(unsigned* p, unsigned v) { const unsigned o = *p; *p = o + v; return o; }
inline unsigned long long atomicAdd(unsigned long long* p, unsigned long long v) {
    const unsigned long long o = *p;
    *p = o + v;
    return o;
}
inline int atomicAdd(int* p, int v) { const int o = *p; *p = o + v; return o; }
inline int atomicOr(int* p, int v) { const int o = *p; *p = o | v; return o; }
inline unsigned long long wall_clock64() { return 0ull; }
inline float __int_as_float(int v) { return simt::from_bits<float>((uint64_t)(uint32_t)v); }
inline int __float_as_int(float v) { return (int)(uint32_t)simt::to_bits(v); }
// double -> float rounded toward -inf / +inf: round to nearest, then step one
// float toward the bound when that rounding went the other way
inline float __double2float_rd(double x) {
    float f = (float)x;
    if ((double)f > x) f = std::nextafter(f, -INFINITY);
    return f;
}
inline float __double2float_ru(double x) {
    float f = (float)x;
    if ((double)f < x) f = std::nextafter(f, INFINITY);
    return f;
}
