// big.hpp -- fused two-step kernel for batches of long 1D complex transforms
// (lengths beyond one work-group's LDS: FP32 2^15..2^16, FP64 2^14..2^16).
//
// N = N1*N2, n = N2*n1 + n2, k = k1 + N1*k2:
//   step 1: for every column n2, Y[k1][n2] = W_N^{n2*k1} * DFT_N1(x[N2*n1 + n2])
//   step 2: for every row k1,    X[k1 + N1*k2] = DFT_N2(Y[k1][.])
// One work-group owns one transform at a time and runs both steps back to
// back through a private scratch slot of N points (written, then re-read by
// the same work-group while it is still in L2 / the MALL), so the transform
// reads and writes HBM once.  Step 1 writes the slot transposed
// (slot[n2*N1 + k1], each column contiguous, staged through LDS), which lets
// step 2 read it and write the output with neighbouring lanes on
// neighbouring rows k1 -- coalesced on both sides without a second staging.
// Work-groups stride over the batch entries (static round robin; entries are
// independent).  The input is read and the output written with non-temporal
// accesses: streamed once, they would otherwise push the scratch slots out
// of the MALL between the two steps (747 x 2^16 FP32: 303 -> 245 us).
#pragma once
#include "fft_core.hpp"

namespace mfx {

struct BigArgs {
    const void* in;
    void* out;
    void* scratch;            // gridDim.x slots of N points
    const void* tw;           // table in the kernel precision
    const double2* tw64;      // FP64 master table
    const long long* in_off;  // entry offsets (points)
    const long long* out_off;
    unsigned count;           // number of entries
};

template <typename T, int L1LOG, int L2LOG, bool INV, int BT, int WPE>
__global__ void __launch_bounds__(BT, WPE) big_kernel(BigArgs A) {
    using C2 = c2<T>;
    constexpr int N1 = 1 << L1LOG, N2 = 1 << L2LOG;
    constexpr int LOGN = L1LOG + L2LOG;
    constexpr unsigned long long NN = 1ull << LOGN;
    constexpr int E = 16;
    constexpr int TPT1 = N1 / E, TPT2 = N2 / E;
    constexpr int CC = BT / TPT1;   // columns per step-1 chunk
    constexpr int CR = BT / TPT2;   // rows per step-2 chunk
    static_assert(N2 % CC == 0 && N1 % CR == 0, "chunking");
    constexpr int SC1 = slot_len(N1), SC2 = slot_len(N2);
    using Core1 = Core<T, L1LOG, E, INV>;
    using Core2 = Core<T, L2LOG, E, INV>;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    C2* lds = reinterpret_cast<C2*>(smem_raw);
    const C2* __restrict__ tw = reinterpret_cast<const C2*>(A.tw);
    const C2* __restrict__ in = reinterpret_cast<const C2*>(A.in);
    C2* __restrict__ out = reinterpret_cast<C2*>(A.out);
    C2* __restrict__ scr = reinterpret_cast<C2*>(A.scratch) + (long long)blockIdx.x * (long long)NN;
    const double2* twN = A.tw64 + NN;

    const int t = threadIdx.x;
    // step 1 mappings: c-fast for the FFT, row for the staged store
    const int c1 = t % CC, tl1 = t / CC;
    const int c1r = t / TPT1, tl1r = t % TPT1;
    // step 2 mapping: c-fast (lanes on neighbouring rows k1)
    const int c2i = t % CR, tl2 = t / CR;

    for (unsigned e = blockIdx.x; e < A.count; e += gridDim.x) {
        const C2* __restrict__ x = in + A.in_off[e];
        C2* __restrict__ y = out + A.out_off[e];
        // ---------------- step 1: columns ----------------
        for (int col0 = 0; col0 < N2; col0 += CC) {
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
            // transposed store: column n2 contiguous in the slot
            {
                const C2* s = lds + c1r * SC1;
                C2* dst = scr + (long long)(col0 + c1r) * N1;
#pragma unroll
                for (int m = 0; m < E; ++m) dst[tl1r + m * TPT1] = s[lpad(tl1r + m * TPT1)];
            }
            __syncthreads();
        }
        // ---------------- step 2: rows ----------------
        for (int row0 = 0; row0 < N1; row0 += CR) {
            const int k1 = row0 + c2i;
            C2 v[E];
#pragma unroll
            for (int m = 0; m < E; ++m) v[m] = scr[(long long)(tl2 + m * TPT2) * N1 + k1];
            C2* slot = lds + c2i * SC2;
            Core2::run(v, slot, tl2, tw);
#pragma unroll
            for (int m = 0; m < E; ++m) st_nt<T>(&y[(long long)(tl2 + m * TPT2) * N1 + k1], v[m]);
            __syncthreads();
        }
    }
}

// LDS bytes a big_kernel instance needs
template <typename T, int L1LOG, int L2LOG, int BT>
constexpr size_t big_lds_bytes() {
    constexpr int N1 = 1 << L1LOG, N2 = 1 << L2LOG;
    constexpr int CC = BT / (N1 / 16), CR = BT / (N2 / 16);
    constexpr size_t a = (size_t)CC * slot_len(N1), b = (size_t)CR * slot_len(N2);
    return (a > b ? a : b) * sizeof(c2<T>);
}

}  // namespace mfx
