// multi.hpp -- one persistent launch for every short 1D complex group of a
// batch (lengths 2^4 .. 2^12 of one direction).
//
// The groups' work is cut into items of C = 256 / TPT transforms (TPT = L/16
// lanes per transform, 16 points per lane: one 256-lane work-group's worth,
// 4096 points, whatever the length), numbered largest length first.  Each
// work-group takes the next item from a device counter (one atomic per item,
// no waiting on other work-groups), finds its group, and runs the tile
// kernel's core for that length: contiguous entries read from and written to
// the owned positions directly, LDS exchanges between the radix-16 passes
// (wavefront-local when a transform fits one wavefront).  Items of one
// launch stream through every CU back to back, so the short groups pay one
// launch and one ramp instead of one per length.  The counter is cleared by
// a one-lane kernel queued in front of each launch.  The same kernel also
// runs one item per work-group (no counter, grid = item count): faster when
// the launch has the GPU to itself, while the persistent form leaves CUs to
// other groups running beside it.
#pragma once
#include "fft_core.hpp"

namespace mfx {

constexpr int kMultiThreads = 256;
constexpr int kMultiMaxGroups = 16;

struct MultiGroup {
    const long long* in_off;   // entry offsets (complex items)
    const long long* out_off;
    unsigned count;            // entries
    unsigned item0;            // first item of the group
    int logL;                  // 4 .. 12
};

struct MultiArgs {
    const void* in;
    void* out;
    const void* tw;            // table in the kernel precision
    unsigned* counter;         // next item (cleared before the launch)
    unsigned total;            // items of all groups
    int ngroups;
    MultiGroup g[kMultiMaxGroups];
};

// the counter reset (a template: instantiated once, where registered)
template <typename U>
__global__ void multi_reset_kernel(U* counter) {
    *counter = U(0);
}

template <typename T, int LOGL, bool INV>
__device__ __forceinline__ void multi_item(const MultiArgs& A, const MultiGroup& G, unsigned item,
                                           c2<T>* lds) {
    using C2 = c2<T>;
    constexpr int L = 1 << LOGL, E = 16, TPT = L / E, C = kMultiThreads / TPT;
    // re-read per item: lane-dependent twiddle reads must not be hoisted
    // out of the item loop (the invariant reads of all nine lengths would
    // pin their registers for the whole kernel)
    const int t = opaque((int)threadIdx.x);
    const int c = t / TPT, tl = t - (t / TPT) * TPT;
    const unsigned idx = item * (unsigned)C + (unsigned)c;
    const bool valid = idx < G.count;
    // an invalid transform reads entry 0 (always readable) and stores nothing
    const long long ib = G.in_off[valid ? idx : 0u], ob = G.out_off[valid ? idx : 0u];
    const C2* __restrict__ x = reinterpret_cast<const C2*>(A.in) + ib;
    C2 v[E];
#pragma unroll
    for (int m = 0; m < E; ++m) v[m] = ld_nt<T>(&x[tl + m * TPT]);  // read once
    Core<T, LOGL, E, INV>::run(v, lds + c * slot_len(L), tl, reinterpret_cast<const C2*>(A.tw), TPT <= 64);
    if (valid) {
        C2* __restrict__ y = reinterpret_cast<C2*>(A.out) + ob;
#pragma unroll
        for (int m = 0; m < E; ++m) st_nt<T>(&y[tl + m * TPT], v[m]);  // written once
    }
}

// item `item` of the launch: find its group, run it
template <typename T, bool INV>
__device__ __forceinline__ void multi_dispatch(const MultiArgs& A, unsigned item, c2<T>* lds) {
    int gi = 0;
    while (gi + 1 < A.ngroups && item >= A.g[gi + 1].item0) ++gi;
    const MultiGroup& G = A.g[gi];
    const unsigned li = item - G.item0;
    switch (G.logL) {
        case 4: multi_item<T, 4, INV>(A, G, li, lds); break;
        case 5: multi_item<T, 5, INV>(A, G, li, lds); break;
        case 6: multi_item<T, 6, INV>(A, G, li, lds); break;
        case 7: multi_item<T, 7, INV>(A, G, li, lds); break;
        case 8: multi_item<T, 8, INV>(A, G, li, lds); break;
        case 9: multi_item<T, 9, INV>(A, G, li, lds); break;
        case 10: multi_item<T, 10, INV>(A, G, li, lds); break;
        case 11: multi_item<T, 11, INV>(A, G, li, lds); break;
        default: multi_item<T, 12, INV>(A, G, li, lds); break;
    }
}

// counter != nullptr: persistent, items claimed from the counter;
// counter == nullptr: one item per work-group (item = work-group index,
// the grid is the item count; no counter, no reset launch)
template <typename T, bool INV>
__global__ void __launch_bounds__(kMultiThreads, sizeof(T) == 8 ? 2 : 4) multi_kernel(MultiArgs A) {
    using C2 = c2<T>;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    C2* lds = reinterpret_cast<C2*>(smem_raw);
    __shared__ unsigned s_item[2];
    if (A.counter == nullptr) {
        if (blockIdx.x < A.total) multi_dispatch<T, INV>(A, blockIdx.x, lds);
        return;
    }
    for (int it = 0;; ++it) {
        // item broadcast through LDS; alternating cells, so that the write of
        // the next item never races with the read of this one
        if (threadIdx.x == 0) s_item[it & 1] = atomicAdd(A.counter, 1u);
        __syncthreads();  // also: every lane is done with the previous item's slots
        const unsigned item = s_item[it & 1];
        if (item >= A.total) break;  // uniform over the work-group
        multi_dispatch<T, INV>(A, item, lds);
    }
}

// LDS of any item: C transforms of slot_len(L) points, largest at L = 2^12
template <typename T>
constexpr size_t multi_lds_bytes() {
    size_t m = 0;
    for (int lg = 4; lg <= 12; ++lg) {
        const size_t b = (size_t)(kMultiThreads / ((1 << lg) / 16)) * slot_len(1 << lg) * sizeof(c2<T>);
        m = b > m ? b : m;
    }
    return m;
}

}  // namespace mfx
