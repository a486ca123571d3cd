// csplit.hpp -- column tile with a radix-2 split along the transformed axis.
// For complex passes whose transforms are adjacent in memory on both sides
// (columns: item stride in_e / out_e along the axis, neighbouring
// transforms one point apart), C transforms per work-group in the c-fast
// lane mapping, direct I/O.  Only half of each transform's L points pass
// through the LDS at a time -- the other half waits in registers -- so a
// work-group covers twice the columns of the whole-transform tile in the
// same LDS: for FP32 L = 512, 32 columns (256-byte row segments, two rows
// per load instruction) with two work-groups per CU instead of 16 columns
// (128-byte segments).  On the 2 MiB-strided outer axis of a 512^3 array
// the wider segments are what the pass is bound by.
//
// Decimation in frequency, as in split2.hpp:
//     a[n] = x[n] + x[n+H],   b[n] = (x[n] - x[n+H]) * W_L^n    (n < H)
//     X[2k] = DFT_H(a)[k],    X[2k+1] = DFT_H(b)[k]
// (inverse: conjugate twiddle, inverse DFT_H).  Lane tl of a transform owns
// n = tl + m*TPT of both halves on input and k = tl + m*TPT on output, so
// its stores go to the neighbouring rows 2k and 2k+1.
#pragma once
#include "tile.hpp"

namespace mfx {

// E: points per lane of the half-length FFTs; WPE: minimum waves per SIMD
// (register budget): 4 lets two 512-lane work-groups share a CU
template <typename T, int LOGH, int E, bool INV, int WPE, bool NT>
__global__ void __launch_bounds__(512, WPE) csplit_kernel(TileArgs A) {
    using C2 = c2<T>;
    constexpr int H = 1 << LOGH, L = 2 * H, TPT = H / E;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    C2* lds = reinterpret_cast<C2*>(smem_raw);
    const C2* __restrict__ tw = reinterpret_cast<const C2*>(A.tw);
    const int t = threadIdx.x;
    const int c = t & (A.C - 1), tl = t >> A.lgC;
    const unsigned tile = tile_id(A);
    const TransformRef me = locate(A, tile * (unsigned)A.C + (unsigned)c);
    C2* slot = lds + (long long)c * A.sc;
    C2 v[E], w[E];
    if (me.valid) {
        const C2* __restrict__ in = reinterpret_cast<const C2*>(A.in) + me.ib + (long long)tl * A.in_e;
        const long long step = A.in_e * TPT, half = A.in_e * H;
#pragma unroll
        for (int m = 0; m < E; ++m) {
            v[m] = NT ? ld_nt<T>(&in[m * step]) : in[m * step];
            w[m] = NT ? ld_nt<T>(&in[m * step + half]) : in[m * step + half];
        }
    } else {
        // (a partial last tile: every lane still reaches the core's barriers)
#pragma unroll
        for (int m = 0; m < E; ++m) v[m] = w[m] = mk<T>(0, 0);
    }
#pragma unroll
    for (int m = 0; m < E; ++m) {
        const C2 e = twl<T>(tw, (unsigned)(L + tl + m * TPT));  // W_L^n
        const C2 s = add<T>(v[m], w[m]);
        const C2 d = sub<T>(v[m], w[m]);
        v[m] = s;
        w[m] = INV ? cmulc<T>(d, e) : cmul<T>(d, e);
    }
    Core<T, LOGH, E, INV>::run(v, slot, tl, tw);  // v: X[2k] at k = tl + m*TPT
    __syncthreads();                              // the slot is free for the second half
    Core<T, LOGH, E, INV>::run(w, slot, tl, tw);  // w: X[2k+1]
    if (me.valid) {
        C2* __restrict__ out = reinterpret_cast<C2*>(A.out) + me.ob;
        const long long es = A.out_e;
#pragma unroll
        for (int m = 0; m < E; ++m) {
            const long long k = tl + m * TPT;
            if constexpr (NT) {
                st_nt<T>(&out[2 * k * es], v[m]);
                st_nt<T>(&out[(2 * k + 1) * es], w[m]);
            } else {
                out[2 * k * es] = v[m];
                out[(2 * k + 1) * es] = w[m];
            }
        }
    }
    if (A.hflag) signal_host(A);
}

// LDS bytes per transform of a csplit_kernel instance (one half-length slot)
template <typename T, int LOGH>
constexpr size_t csplit_slot_bytes() {
    return (size_t)slot_len(1 << LOGH) * sizeof(c2<T>);
}

}  // namespace mfx
