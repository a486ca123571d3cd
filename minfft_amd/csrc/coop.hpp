// coop.hpp -- XCD-cooperative two-step kernel for batches of long 1D complex
// transforms, keeping the intermediate in the XCD's L2.
//
// STATUS: disabled, unreachable from the engine.  Both GPU runs of it (FP64,
// N = 2^14, 300 entries) hung -- the first with device-scope load polling,
// the second with the RMW polling and bounded waits below -- and the cause
// is not established.  Kept for study; do not enable without one.
//
// Same factorisation as big.hpp (N = N1*N2; step 1 = N2 columns of length N1
// with the W_N^{n2*k1} twiddle, written transposed; step 2 = N1 rows of
// length N2), but the work of one transform is spread over many work-groups
// of ONE XCD, so that only a few transforms per XCD are in flight and their
// intermediates (a few MiB) stay in that XCD's 4 MiB L2 instead of making a
// round trip to HBM.
//
// Scheduling.  Each work-group reads its XCD number from the hardware
// (XCC_ID) and pulls work items from that XCD's queue (an atomic counter).
// Entries e with e % 8 == xcd belong to XCD xcd; their local index g orders
// the queue in rounds r = 0, 1, ...:
//     round r = [CH1 step-1 chunks of transform r] [CH2 step-2 chunks of
//                transform r - LA]
// A step-2 chunk of g waits until all CH1 step-1 chunks of g are done; a
// step-1 chunk of g waits until the scratch slot (g mod S) has been released
// by all CH2 step-2 chunks of g - S.  With S > LA every dependency is on an
// item earlier in the same queue, i.e. already claimed by a running
// work-group, so the waits always end, however many work-groups are resident.
//
// Coherence.  All work-groups sharing a slot sit on the same XCD and share
// its L2.  Counters are only ever touched by read-modify-write atomics
// (claims, increments, and polls as "add 0"), which all execute at the same
// place; device-scope atomic loads would bypass the cache and miss
// increments that live in L2.  A producer waits for its scratch stores to be
// acknowledged by L2 (s_waitcnt 0), synchronises the work-group and then
// increments the chunk counter.  A consumer polls, synchronises, invalidates
// its CU's vector L1 (buffer_inv sc0: a slot's old lines may still be cached
// there from an earlier transform) and reads the slot with ordinary loads,
// which hit the L2 lines the producers wrote.
// Every wait is bounded: after kSpinLimit polls a work-group raises a shared
// abort flag and all waits return at once, so a scheduling or coherence
// failure ends the kernel (with wrong results) instead of hanging it.
#pragma once
#include "fft_core.hpp"

namespace mfx {

constexpr int kXcds = 8;

struct CoopArgs {
    const void* in;
    void* out;
    void* scratch;            // kXcds * slots slots of N points
    const void* tw;
    const double2* tw64;
    const long long* in_off;  // entry offsets (points)
    const long long* out_off;
    unsigned* ctrs;           // [kXcds heads][count done1][count done2][abort], zeroed per launch
    unsigned count;
    int slots;                // S
    int lookahead;            // LA < S
};

constexpr unsigned kSpinLimit = 1u << 22;

// current value of a counter, read with an atomic RMW (add of a zero the
// compiler cannot see) so that it executes where the increments do
__device__ __forceinline__ unsigned rmw_read(unsigned* c) {
    return __hip_atomic_fetch_add(c, (unsigned)opaque(0), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// wait until *c >= target; false if the shared abort flag is (or gets) set
__device__ __forceinline__ bool wait_count(unsigned* c, unsigned target, unsigned* abort_flag) {
    for (unsigned it = 0;; ++it) {
        if (rmw_read(c) >= target) return true;
        if (rmw_read(abort_flag) != 0) return false;
        if (it >= kSpinLimit) {
            __hip_atomic_fetch_add(abort_flag, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return false;
        }
        __builtin_amdgcn_s_sleep(1);
    }
}

// invalidate this CU's vector L1 so later loads are served by L2
__device__ __forceinline__ void l1_invalidate() { asm volatile("buffer_inv sc0" ::: "memory"); }

static __global__ void coop_reset(unsigned* ctrs, unsigned n) {
    const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) ctrs[i] = 0;
}

template <typename T, int L1LOG, int L2LOG, bool INV, int BT, int WPE>
__global__ void __launch_bounds__(BT, WPE) coop_kernel(CoopArgs A) {
    using C2 = c2<T>;
    constexpr int N1 = 1 << L1LOG, N2 = 1 << L2LOG;
    constexpr int LOGN = L1LOG + L2LOG;
    constexpr unsigned long long NN = 1ull << LOGN;
    constexpr int E = 16;
    constexpr int TPT1 = N1 / E, TPT2 = N2 / E;
    constexpr int CC = BT / TPT1;   // columns per step-1 chunk
    constexpr int CR = BT / TPT2;   // rows per step-2 chunk
    constexpr int CH1 = N2 / CC, CH2 = N1 / CR;
    static_assert(N2 % CC == 0 && N1 % CR == 0, "chunking");
    constexpr int SC1 = slot_len(N1), SC2 = slot_len(N2);
    using Core1 = Core<T, L1LOG, E, INV>;
    using Core2 = Core<T, L2LOG, E, INV>;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ unsigned s_item;
    C2* lds = reinterpret_cast<C2*>(smem_raw);
    const C2* __restrict__ tw = reinterpret_cast<const C2*>(A.tw);
    const C2* __restrict__ in = reinterpret_cast<const C2*>(A.in);
    C2* __restrict__ out = reinterpret_cast<C2*>(A.out);
    const double2* twN = A.tw64 + NN;

    // the XCD this work-group runs on (hardware register XCC_ID, bits 3:0)
    const unsigned xcd = (unsigned)__builtin_amdgcn_s_getreg((3 << 11) | 20) % kXcds;
    unsigned* head = A.ctrs + xcd;
    unsigned* done1 = A.ctrs + kXcds;
    unsigned* done2 = A.ctrs + kXcds + A.count;
    unsigned* abort_flag = A.ctrs + kXcds + 2 * A.count;
    const unsigned G = (A.count > xcd) ? (A.count - xcd + kXcds - 1) / kXcds : 0;  // this XCD's transforms
    const int S = A.slots, LA = A.lookahead;
    const unsigned per_round = CH1 + CH2;
    const unsigned total = (G + (unsigned)LA) * per_round;

    const int t = threadIdx.x;
    const int c1 = t % CC, tl1 = t / CC;
    const int c1r = t / TPT1, tl1r = t % TPT1;
    const int c2i = t % CR, tl2 = t / CR;

    for (;;) {
        if (t == 0) s_item = atomicAdd(head, 1u);
        __syncthreads();
        const unsigned item = s_item;
        __syncthreads();
        if (item >= total) break;
        const unsigned r = item / per_round, j = item % per_round;
        if (j < (unsigned)CH1) {
            // ---------------- step 1 chunk j of transform r ----------------
            const unsigned g = r;
            if (g >= G) continue;
            const unsigned e = g * kXcds + xcd;
            C2* scr = reinterpret_cast<C2*>(A.scratch) + ((long long)xcd * S + (long long)(g % S)) * (long long)NN;
            if (g >= (unsigned)S) {
                if (t == 0) (void)wait_count(done2 + (e - (unsigned)S * kXcds), CH2, abort_flag);
                __syncthreads();
            }
            const C2* __restrict__ x = in + A.in_off[e];
            const int col0 = (int)j * CC;
            const int n2 = col0 + c1;
            C2 v[E];
#pragma unroll
            for (int m = 0; m < E; ++m) v[m] = x[(long long)(tl1 + m * TPT1) * N2 + n2];
            C2* slot = lds + c1 * SC1;
            Core1::run(v, slot, tl1, tw);
            double2 w = twN[((unsigned long long)n2 * (unsigned)tl1) & (NN - 1)];
            const double2 st = twN[((unsigned long long)n2 * (unsigned)TPT1) & (NN - 1)];
#pragma unroll
            for (int m = 0; m < E; ++m) {
                const C2 wt = mk<T>((T)w.x, (T)w.y);
                v[m] = INV ? cmulc<T>(v[m], wt) : cmul<T>(v[m], wt);
                if (m + 1 < E) w = make_double2(w.x * st.x - w.y * st.y, w.x * st.y + w.y * st.x);
            }
            Core1::write_owned(v, slot, tl1);
            __syncthreads();
            {
                const C2* s = lds + c1r * SC1;
                C2* dst = scr + (long long)(col0 + c1r) * N1;
#pragma unroll
                for (int m = 0; m < E; ++m) dst[tl1r + m * TPT1] = s[lpad(tl1r + m * TPT1)];
            }
            __builtin_amdgcn_s_waitcnt(0);  // stores acknowledged by L2
            __syncthreads();
            if (t == 0) __hip_atomic_fetch_add(done1 + e, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            // ---------------- step 2 chunk of transform r - LA ----------------
            if (r < (unsigned)LA) continue;
            const unsigned g = r - (unsigned)LA;
            if (g >= G) continue;
            const unsigned e = g * kXcds + xcd;
            const C2* scr = reinterpret_cast<const C2*>(A.scratch) +
                            ((long long)xcd * S + (long long)(g % S)) * (long long)NN;
            if (t == 0) (void)wait_count(done1 + e, CH1, abort_flag);
            __syncthreads();
            l1_invalidate();
            C2* __restrict__ y = out + A.out_off[e];
            const int k1 = (int)(j - CH1) * CR + c2i;
            C2 v[E];
#pragma unroll
            for (int m = 0; m < E; ++m) v[m] = scr[(long long)(tl2 + m * TPT2) * N1 + k1];
            C2* slot = lds + c2i * SC2;
            Core2::run(v, slot, tl2, tw);
#pragma unroll
            for (int m = 0; m < E; ++m) y[(long long)(tl2 + m * TPT2) * N1 + k1] = v[m];
            __syncthreads();
            if (t == 0) __hip_atomic_fetch_add(done2 + e, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

template <typename T, int L1LOG, int L2LOG, int BT>
constexpr size_t coop_lds_bytes() {
    constexpr int N1 = 1 << L1LOG, N2 = 1 << L2LOG;
    constexpr int CC = BT / (N1 / 16), CR = BT / (N2 / 16);
    constexpr size_t a = (size_t)CC * slot_len(N1), b = (size_t)CR * slot_len(N2);
    return (a > b ? a : b) * sizeof(c2<T>);
}

}  // namespace mfx
