// zdx_sort.h — the introspective sort whose element permutation equals that of
// GNU libstdc++'s std::sort (GCC 11): median-of-three pivot moved to the
// front, unguarded Hoare partition, recursion on the right part and loop on
// the left, depth limit 2*floor(log2 n) falling back to heap sort, then a
// final insertion sort with threshold 16.
//
// Why it exists: the reference orders loop start vertices, contact components
// and T-junction splits with std::sort (src/main.cpp:897, 945, 977), which is
// not stable; for parity of loop topology the GPU must permute equal keys the
// same way.  The steps follow the published introsort algorithm (Musser 1997)
// as libstdc++ configures it; the code is this project's own.
//
// Works on a caller-provided array of trivially-copyable elements through an
// accessor (so it can run on strided per-thread scratch on the GPU) and is
// __host__ __device__ so tests compare it with std::sort on the CPU.
#pragma once
#include <stdint.h>

#ifndef ZDX_HD
#define ZDX_HD __host__ __device__
#endif

namespace zdx {

// Stack of the introsort's pending right-hand ranges (first, last, depth).
#if defined(__HIP_DEVICE_COMPILE__)
// On the GPU every lane of the wave runs the sort with identical values, so
// entry k is kept in lane k's registers (64 entries per wave; the depth limit
// keeps the stack below that) and read back with a lane read, instead of a
// per-lane array in scratch memory.
struct IntroStack {
    int f, l, d;   // this lane's entry
    __device__ static int lane() { return (int)__builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }
    __device__ void put(int k, int vf, int vl, int vd) {
        if (lane() == k) { f = vf; l = vl; d = vd; }
    }
    __device__ void get(int k, int& vf, int& vl, int& vd) const {
        const int ks = __builtin_amdgcn_readfirstlane(k);
        vf = __builtin_amdgcn_readlane(f, ks);
        vl = __builtin_amdgcn_readlane(l, ks);
        vd = __builtin_amdgcn_readlane(d, ks);
    }
};
#else
struct IntroStack {
    int f[64], l[64], d[64];
    void put(int k, int vf, int vl, int vd) { f[k] = vf; l[k] = vl; d[k] = vd; }
    void get(int k, int& vf, int& vl, int& vd) const { vf = f[k]; vl = l[k]; vd = d[k]; }
};
#endif

template <class Arr, class Less>
struct IntroSorter {
    using Stack = IntroStack;
    Arr a;        // a[i] read/write, element type Arr::value_type
    Less less;    // strict weak ordering on values
    using T = typename Arr::value_type;

    ZDX_HD void swap_at(int i, int j) { T t = a.get(i); a.set(i, a.get(j)); a.set(j, t); }

    ZDX_HD static int floor_log2(int n) { int k = 0; while (n > 1) { n >>= 1; ++k; } return k; }

    // median of (x, y, z) moved to position r
    ZDX_HD void median_to_first(int r, int x, int y, int z) {
        if (less(a.get(x), a.get(y))) {
            if (less(a.get(y), a.get(z))) swap_at(r, y);
            else if (less(a.get(x), a.get(z))) swap_at(r, z);
            else swap_at(r, x);
        } else if (less(a.get(x), a.get(z))) swap_at(r, x);
        else if (less(a.get(y), a.get(z))) swap_at(r, z);
        else swap_at(r, y);
    }

    // partition [first, last) around the pivot value at index pv
    ZDX_HD int unguarded_partition(int first, int last, int pv) {
        while (true) {
            while (less(a.get(first), a.get(pv))) ++first;
            --last;
            while (less(a.get(pv), a.get(last))) --last;
            if (!(first < last)) return first;
            swap_at(first, last);
            ++first;
        }
    }

    ZDX_HD int partition_pivot(int first, int last) {
        int mid = first + (last - first) / 2;
        median_to_first(first, first + 1, mid, last - 1);
        return unguarded_partition(first + 1, last, first);
    }

    // heap helpers on the range [first, first+len)
    ZDX_HD void push_heap(int first, int hole, int top, T v) {
        int parent = (hole - 1) / 2;
        while (hole > top && less(a.get(first + parent), v)) {
            a.set(first + hole, a.get(first + parent));
            hole = parent;
            parent = (hole - 1) / 2;
        }
        a.set(first + hole, v);
    }
    ZDX_HD void adjust_heap(int first, int hole, int len, T v) {
        const int top = hole;
        int child = hole;
        while (child < (len - 1) / 2) {
            child = 2 * (child + 1);
            if (less(a.get(first + child), a.get(first + (child - 1)))) child--;
            a.set(first + hole, a.get(first + child));
            hole = child;
        }
        if ((len & 1) == 0 && child == (len - 2) / 2) {
            child = 2 * (child + 1);
            a.set(first + hole, a.get(first + (child - 1)));
            hole = child - 1;
        }
        push_heap(first, hole, top, v);
    }
    ZDX_HD void make_heap(int first, int last) {
        int len = last - first;
        if (len < 2) return;
        int parent = (len - 2) / 2;
        while (true) {
            T v = a.get(first + parent);
            adjust_heap(first, parent, len, v);
            if (parent == 0) return;
            parent--;
        }
    }
    ZDX_HD void pop_heap(int first, int last, int result) {
        T v = a.get(result);
        a.set(result, a.get(first));
        adjust_heap(first, 0, last - first, v);
    }
    ZDX_HD void heap_sort(int first, int last) {
        make_heap(first, last);
        while (last - first > 1) {
            --last;
            pop_heap(first, last, last);
        }
    }

    ZDX_HD void introsort_loop(int first, int last, int depth) {
        if (last - first <= 16) return;   // left to the final insertion sort
        // explicit stack for the right-hand recursion (depth <= 2*log2(n) < 64)
        Stack st;
        int sp = 0;
        st.put(sp, first, last, depth); ++sp;
        while (sp > 0) {
            --sp;
            int f, l, d;
            st.get(sp, f, l, d);
            while (l - f > 16) {
                if (d == 0) { heap_sort(f, l); break; }
                --d;
                int cut = partition_pivot(f, l);
                // the reference recurses on [cut, l) before continuing with [f, cut)
                // order of processing is irrelevant to the result: disjoint ranges
                st.put(sp, cut, l, d); ++sp;
                l = cut;
            }
        }
    }

    ZDX_HD void linear_insert(int last) {
        T v = a.get(last);
        int next = last - 1;
        while (less(v, a.get(next))) {
            a.set(last, a.get(next));
            last = next;
            --next;
        }
        a.set(last, v);
    }
    ZDX_HD void insertion_sort(int first, int last) {
        if (first == last) return;
        for (int i = first + 1; i != last; ++i) {
            if (less(a.get(i), a.get(first))) {
                T v = a.get(i);
                for (int k = i; k > first; --k) a.set(k, a.get(k - 1));
                a.set(first, v);
            } else {
                linear_insert(i);
            }
        }
    }
    ZDX_HD void final_insertion_sort(int first, int last) {
        if (last - first > 16) {
            insertion_sort(first, first + 16);
            for (int i = first + 16; i != last; ++i) linear_insert(i);
        } else {
            insertion_sort(first, last);
        }
    }

    ZDX_HD void sort(int n) {
        if (n <= 1) return;
        introsort_loop(0, n, floor_log2(n) * 2);
        final_insertion_sort(0, n);
    }
};

template <class Arr, class Less>
ZDX_HD inline void ref_sort(Arr a, int n, Less less) {
    IntroSorter<Arr, Less> s{a, less};
    s.sort(n);
}

}  // namespace zdx
