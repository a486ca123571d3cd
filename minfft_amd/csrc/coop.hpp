// coop.hpp -- XCD-cooperative two-step kernel for batches of long 1D complex
// transforms, with no work-group ever waiting on another.
//
// Same factorisation as big.hpp (N = N1*N2; step 1 = N2 columns of length N1
// with the W_N^{n2*k1} twiddle, written transposed into the transform's
// scratch slot; step 2 = N1 rows of length N2, written to the output), but
// the CH1 step-1 chunks and CH2 step-2 chunks of one transform are spread
// over many work-groups of ONE XCD.  Few transforms per XCD are then in
// flight, and their intermediates (512 KiB each at N = 2^16 FP32) stay in
// that XCD's 4 MiB L2 between the steps instead of making a round trip
// through the MALL / HBM, as the fused kernel's one-slot-per-work-group
// layout (256 MiB of slots in flight) does.
//
// Work.  Each work-group reads its XCD number from the hardware register
// XCC_ID and takes items from that XCD's queue (one atomic counter).  Entry
// e belongs to queue e % nq (local index g = e / nq); the queue interleaves
//     round r = [CH1 step-1 chunks of transform r][CH2 step-2 chunks of
//                transform r - LA]
// so step 2 of a transform is claimed LA rounds after its step 1.  Every
// entry has its own slot (no slot is reused inside a launch), so step-1
// chunks depend on nothing.
//
// Dependency without waiting.  Each transform has one 32-bit word: bits 0-7
// count its finished step-1 chunks, bit 8+c marks step-2 chunk c as
// deferred.  A work-group that claims step-2 chunk c sets bit 8+c with an
// atomic OR and looks at the old word: if all CH1 step-1 chunks are done it
// runs the chunk, otherwise it leaves it (deferred) and claims its next item.
// A work-group that finishes a step-1 chunk adds 1 with an atomic ADD; the
// one whose ADD completes step 1 runs every chunk whose bit was already set.
// The two atomics on one word are ordered, so each chunk runs exactly once
// (by its claimant if its OR came after the last ADD, else by the finisher),
// and since nothing waits the kernel always drains: every loop iteration
// claims a new queue position, and the queue is finite.
//
// Cost.  Every item carries up to three atomic round trips on its work-
// group's critical path (the claim, the wait for the slot stores'
// acknowledgement, the ADD), about 2-5 us against 9-12 us of work for a
// 64 KiB FP32 chunk.  Measured on MI355X: FP64 2^16 x 300 entries 184 us
// against 300 us for the fused kernel (whose 1 MiB-per-work-group slots
// overflow the MALL), FP32 2^16 x 747 entries 243-279 us against 215 us
// (its 256 MiB of slots are served by the MALL).  So the kernel is the
// default for FP64 and a tuner option for FP32.  (Claiming the next item
// ahead and releasing a chunk one item late, to hide the round trips,
// measured slower: 350 us FP32.)
//
// Coherence.  All work-groups that touch a slot run on the same XCD and share
// its L2; the protocol is that of work-groups split over CUs: a producer
// waits until its slot stores are acknowledged by L2 (s_waitcnt vmcnt(0)),
// synchronises the work-group, then does its ADD; a consumer learns from an
// atomic that the chunks are done, synchronises, invalidates its CU's vector
// L1 (buffer_inv sc0) and reads the slot with ordinary loads, served by L2.
// Counters are only touched by read-modify-write atomics (one place of
// execution for every access); a reset kernel zeroes them before each launch.
#pragma once
#include "fft_core.hpp"

namespace mfx {

struct CoopArgs {
    const void* in;
    void* out;
    void* scratch;             // two-step scratch: one slot of N points per entry
    const void* tw;            // table in the kernel precision
    const double2* tw64;       // FP64 master table
    const long long* in_off;   // entry offsets (points)
    const long long* out_off;
    const long long* slot_off;  // per-entry slot offsets into scratch (points)
    unsigned* ctr;             // [nq] queue heads, then [count] per-entry words
    unsigned count;            // entries
    unsigned nq;               // queues (XCDs of the device, 1..8)
    unsigned lookahead;        // LA: rounds between a transform's step 1 and step 2
};

__global__ void coop_reset_kernel(unsigned* ctr, unsigned n);

// XCD of the CU running this wavefront (hardware register XCC_ID, bits 3:0)
__device__ __forceinline__ unsigned xcc_id() { return (unsigned)__builtin_amdgcn_s_getreg((3 << 11) | 20) & 15u; }

__device__ __forceinline__ unsigned atom_add(unsigned* p, unsigned v) {
    return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned atom_or(unsigned* p, unsigned v) {
    return __hip_atomic_fetch_or(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
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
    static_assert(CH1 < 256 && CH2 <= 24, "per-entry word layout");
    constexpr int SC1 = slot_len(N1), SC2 = slot_len(N2);
    using Core1 = Core<T, L1LOG, E, INV>;
    using Core2 = Core<T, L2LOG, E, INV>;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ unsigned s_cmd[2];
    C2* lds = reinterpret_cast<C2*>(smem_raw);
    const C2* __restrict__ tw = reinterpret_cast<const C2*>(A.tw);
    const double2* twN = A.tw64 + NN;

    const unsigned q = xcc_id() % A.nq;
    unsigned* head = A.ctr + q;
    unsigned* word = A.ctr + A.nq;
    const unsigned G = A.count > q ? (A.count - q + A.nq - 1) / A.nq : 0;  // this queue's entries
    const unsigned LA = A.lookahead;
    constexpr unsigned per_round = CH1 + CH2;
    const unsigned total = (G + LA) * per_round;

    const int t = threadIdx.x;
    const int c1 = t % CC, tl1 = t / CC;
    const int c1r = t / TPT1, tl1r = t % TPT1;
    const int c2i = t % CR, tl2 = t / CR;

    // step-2 chunk c of entry e: rows k1 = c*CR .. c*CR+CR-1
    auto step2 = [&](unsigned e, int c) {
        asm volatile("buffer_inv sc0" ::: "memory");  // slot lines come from L2
        const C2* __restrict__ scr = reinterpret_cast<const C2*>(A.scratch) + A.slot_off[e];
        C2* __restrict__ y = reinterpret_cast<C2*>(A.out) + A.out_off[e];
        const int k1 = c * CR + c2i;
        C2 v[E];
#pragma unroll
        for (int m = 0; m < E; ++m) v[m] = scr[(long long)(tl2 + m * TPT2) * N1 + k1];
        C2* slot = lds + c2i * SC2;
        Core2::run(v, slot, tl2, tw);
#pragma unroll
        for (int m = 0; m < E; ++m) st_nt<T>(&y[(long long)(tl2 + m * TPT2) * N1 + k1], v[m]);
        __syncthreads();  // LDS free for the next item
    };

    for (;;) {
        // ---- claim the next queue position (thread 0), broadcast the command:
        // s_cmd[0] = 0 done, 1 skip, 2 step 1, 3 step 2; s_cmd[1] = entry*256 + chunk
        if (t == 0) {
            const unsigned s = atom_add(head, 1u);
            unsigned cmd = 1, arg = 0;
            if (s >= total) {
                cmd = 0;
            } else {
                const unsigned r = s / per_round, j = s % per_round;
                if (j < (unsigned)CH1) {
                    if (r < G) { cmd = 2; arg = ((r * A.nq + q) << 8) | j; }
                } else if (r >= LA && r - LA < G) {
                    const unsigned e = (r - LA) * A.nq + q;
                    const unsigned c = j - CH1;
                    const unsigned old = atom_or(word + e, 1u << (8 + c));
                    if ((old & 0xffu) == (unsigned)CH1) { cmd = 3; arg = (e << 8) | c; }
                    // else deferred: the work-group finishing step 1 of e runs it
                }
            }
            s_cmd[0] = cmd;
            s_cmd[1] = arg;
        }
        __syncthreads();
        const unsigned cmd = s_cmd[0], arg = s_cmd[1];
        __syncthreads();
        if (cmd == 0) break;
        if (cmd == 1) continue;
        const unsigned e = arg >> 8;
        if (cmd == 3) {
            step2(e, (int)(arg & 0xffu));
            continue;
        }
        // ---------------- step-1 chunk: columns col0 .. col0+CC-1 ----------------
        const int col0 = (int)(arg & 0xffu) * CC;
        const C2* __restrict__ x = reinterpret_cast<const C2*>(A.in) + A.in_off[e];
        C2* __restrict__ scr = reinterpret_cast<C2*>(A.scratch) + A.slot_off[e];
        const int n2 = col0 + c1;
        C2 v[E];
#pragma unroll
        for (int m = 0; m < E; ++m) v[m] = ld_nt<T>(&x[(long long)(tl1 + m * TPT1) * N2 + n2]);
        C2* slot = lds + c1 * SC1;
        Core1::run(v, slot, tl1, tw);
        // W_N^{n2*k1}, k1 = tl1 + m*TPT1, by an FP64 recurrence over m
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
            // transposed store: column n2 contiguous in the slot
            const C2* sl = lds + c1r * SC1;
            C2* dst = scr + (long long)(col0 + c1r) * N1;
#pragma unroll
            for (int m = 0; m < E; ++m) dst[tl1r + m * TPT1] = sl[lpad(tl1r + m * TPT1)];
        }
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");  // slot stores acknowledged by L2
        __syncthreads();
        if (t == 0) {
            const unsigned old = atom_add(word + e, 1u);
            s_cmd[0] = ((old & 0xffu) + 1u == (unsigned)CH1) ? (old >> 8) : 0u;  // deferred chunks
        }
        __syncthreads();
        unsigned mask = s_cmd[0];
        __syncthreads();
        while (mask) {
            const int c = __builtin_ctz(mask);
            mask &= mask - 1;
            step2(e, c);
        }
    }
}

// LDS bytes a coop_kernel instance needs (besides its two static words)
template <typename T, int L1LOG, int L2LOG, int BT>
constexpr size_t coop_lds_bytes() {
    constexpr int N1 = 1 << L1LOG, N2 = 1 << L2LOG;
    constexpr int CC = BT / (N1 / 16), CR = BT / (N2 / 16);
    constexpr size_t a = (size_t)CC * slot_len(N1), b = (size_t)CR * slot_len(N2);
    return (a > b ? a : b) * sizeof(c2<T>);
}

}  // namespace mfx
