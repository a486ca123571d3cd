// split2.hpp -- single-pass complex FFT of length L = 2H, one work-group
// per transform, whose LDS holds only H points.  For the first length beyond
// one work-group's LDS (FP32 L = 2^15, FP64 L = 2^14) it gives one HBM read
// and one HBM write per point instead of the two of the two-step
// decomposition; at the LDS limit itself (and one below) it halves the LDS
// of a work-group, so that two share a CU and overlap memory and compute.
//
// Radix-2 decimation in frequency:
//     a[n] = x[n] + x[n+H],   b[n] = (x[n] - x[n+H]) * W_L^n    (n < H)
//     X[2k] = DFT_H(a)[k],    X[2k+1] = DFT_H(b)[k]
// (inverse: conjugate twiddle, inverse DFT_H).  Each lane loads x at its
// owned positions n = tl + m*TPT of both halves, forms a and b in
// registers, runs DFT_H(a) through the LDS slot while b waits in registers,
// then DFT_H(b) while DFT_H(a)'s result waits, and writes the pairs
// (X[2k], X[2k+1]) of its owned k -- adjacent lanes, adjacent pairs, one
// 16-byte store per pair in FP32.
// NT: input and output move with non-temporal accesses -- for whole 1D
// transforms (streamed once: in a batch they would otherwise displace the
// fused kernel's scratch slots of the groups beside them) and passes larger
// than the MALL; a row pass of a smaller multi-dimensional array keeps its
// output in the MALL for the next pass.
// Contiguous transforms only (item stride 1 on both sides).
#pragma once
#include "tile.hpp"

namespace mfx {

// E: points per lane of the half-length FFTs (16, or 32 for radix-32
// passes); WPE: minimum waves per SIMD (register budget): 4 lets two
// 512-lane work-groups share a CU when their LDS fits twice
template <typename T, int LOGH, int E, bool INV, int WPE, bool NT>
__global__ void __launch_bounds__((1 << LOGH) / E, WPE) split2_kernel(TileArgs A) {
    using C2 = c2<T>;
    constexpr int H = 1 << LOGH, L = 2 * H, TPT = H / E;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    C2* slot = reinterpret_cast<C2*>(smem_raw);
    const C2* __restrict__ tw = reinterpret_cast<const C2*>(A.tw);
    const int tl = threadIdx.x;
    // one transform per work-group; the grid is exactly the transform count
    const TransformRef me = locate(A, blockIdx.x);
    if (!me.valid) return;  // uniform over the work-group
    const C2* __restrict__ in = reinterpret_cast<const C2*>(A.in) + me.ib;
    C2 v[E], w[E];
#pragma unroll
    for (int m = 0; m < E; ++m) {
        v[m] = NT ? ld_nt<T>(&in[tl + m * TPT]) : in[tl + m * TPT];
        w[m] = NT ? ld_nt<T>(&in[tl + m * TPT + H]) : in[tl + m * TPT + H];
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
    __syncthreads();                              // slot free for the second half
    Core<T, LOGH, E, INV>::run(w, slot, tl, tw);  // w: X[2k+1]
    C2* __restrict__ out = reinterpret_cast<C2*>(A.out) + me.ob;
    // the pair (X[2k], X[2k+1]) as one 16-byte store when it is aligned
    // (FP32, outputs at a 16-byte boundary -- uniform over the work-group)
    if constexpr (sizeof(T) == 4) {
        if ((reinterpret_cast<unsigned long long>(out) & 15) == 0) {
            typedef float V4 __attribute__((ext_vector_type(4)));
#pragma unroll
            for (int m = 0; m < E; ++m) {
                const int k = tl + m * TPT;
                V4 q;
                q.x = v[m].x;
                q.y = v[m].y;
                q.z = w[m].x;
                q.w = w[m].y;
                V4* p = reinterpret_cast<V4*>(out + 2 * k);
                if constexpr (NT) __builtin_nontemporal_store(q, p);
                else *p = q;
            }
            return;
        }
    }
#pragma unroll
    for (int m = 0; m < E; ++m) {
        const int k = tl + m * TPT;
        if constexpr (NT) {
            st_nt<T>(&out[2 * k], v[m]);
            st_nt<T>(&out[2 * k + 1], w[m]);
        } else {
            out[2 * k] = v[m];
            out[2 * k + 1] = w[m];
        }
    }
}

template <typename T, int LOGH>
constexpr size_t split2_lds_bytes() {
    return (size_t)slot_len(1 << LOGH) * sizeof(c2<T>);
}

}  // namespace mfx
