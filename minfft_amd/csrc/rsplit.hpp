// rsplit.hpp -- single-pass complex FFT of a length N = P*M beyond one
// work-group's LDS, spread over P/B work-groups that need no communication.
//
// Radix-P decimation in frequency, one residue class of outputs per FFT:
//     y_q[n]      = W_N^{n q} * sum_{j<P} x[n + j M] * W_P^{j q}     (n < M)
//     X[P k + q]  = DFT_M(y_q)[k]                                   (k < M)
// (inverse: conjugate twiddles, inverse DFT_M).  Work-group (t, r) of
// transform t computes the B residues q = r*B .. r*B+B-1: every lane loads
// x at its owned positions n = tl + m*TPT of all P blocks (contiguous,
// coalesced), forms the B sums in registers, applies the FP64-generated
// twiddle W_N^{n q} (W_N^{q tl} * W_N^{q m TPT}, multiplied in FP64), runs
// DFT_M through the LDS slot (B > 1: the later residues wait in registers
// while the slot serves the earlier ones) and writes X[P k + q] -- B
// adjacent outputs per k, so one B*8-byte store per lane and k.
//
// HBM traffic: the P work-groups of a transform read the same input; the
// grid places them on one XCD, dispatched back to back (work-group g runs on
// XCD g mod 8), so all but the first read of a line hit that XCD's L2 and
// the transform reads HBM about once and writes it once, where a two-step
// decomposition reads and writes it twice.  The LDS slot holds M points,
// which keeps two or more work-groups per CU.
//
// Contiguous transforms only (item stride 1 on both sides).
#pragma once
#include "tile.hpp"

namespace mfx {

// c2 at an unsigned item offset from a uniform base (32-bit vector offset)
template <typename T>
__device__ __forceinline__ c2<T> ld_at(const c2<T>* __restrict__ base, unsigned i) {
    return *reinterpret_cast<const c2<T>*>(reinterpret_cast<const char*>(base) + i * (unsigned)sizeof(c2<T>));
}
template <typename T>
__device__ __forceinline__ void st_at(c2<T>* __restrict__ base, unsigned i, c2<T> v) {
    *reinterpret_cast<c2<T>*>(reinterpret_cast<char*>(base) + i * (unsigned)sizeof(c2<T>)) = v;
}

// Residues q = R*B + b (b < B) of the radix-P split, with compile-time
// rotations W_P^{j q} (multiples of 360/P degrees: exact shortcuts for
// P <= 8), followed by the twiddle W_N^{n q}.
template <typename T, int LOGM, int LOGP, int B, int E, bool INV, int R>
__device__ __forceinline__ void rsplit_combine(c2<T> (&v)[B][E], const c2<T>* __restrict__ in, int tl,
                                               const double2* __restrict__ twN) {
    using C2 = c2<T>;
    constexpr int M = 1 << LOGM, P = 1 << LOGP, TPT = M / E;
#pragma unroll
    for (int j = 0; j < P; ++j) {
        C2 x[E];
#pragma unroll
        for (int m = 0; m < E; ++m) x[m] = ld_at<T>(in, (unsigned)(tl + m * TPT + j * M));
#pragma unroll
        for (int b = 0; b < B; ++b) {
            const int q = R * B + b;
            const int k = (32 / P) * ((j * q) & (P - 1));  // W_P^{jq} = W_32^k
#pragma unroll
            for (int m = 0; m < E; ++m) {
                const C2 r = rot32<T, INV>(x[m], k);
                v[b][m] = (j == 0) ? r : add<T>(v[b][m], r);
            }
        }
    }
#pragma unroll
    for (int b = 0; b < B; ++b) {
        const int q = R * B + b;
        if (q == 0) continue;
        // W_N^{q n}, n = tl + m*TPT: W_N^{q tl} (per lane) times the uniform
        // W_N^{q m TPT}; q*n < N, so no index wraps
        const double2 w0 = twN[(unsigned)(q * tl)];
#pragma unroll
        for (int m = 0; m < E; ++m) {
            const double2 s = twN[(unsigned)(q * m * TPT)];
            const double wx = w0.x * s.x - w0.y * s.y, wy = w0.x * s.y + w0.y * s.x;
            const C2 w = mk<T>((T)wx, (T)wy);
            v[b][m] = INV ? cmulc<T>(v[b][m], w) : cmul<T>(v[b][m], w);
        }
    }
}

// Uniform dispatch on the residue group r (< G) to the specialised combine.
template <typename T, int LOGM, int LOGP, int B, int E, bool INV, int R, int G>
struct RsplitDispatch {
    __device__ static __forceinline__ void go(int r, c2<T> (&v)[B][E], const c2<T>* __restrict__ in, int tl,
                                              const double2* __restrict__ twN) {
        if constexpr (R + 1 == G) {
            rsplit_combine<T, LOGM, LOGP, B, E, INV, R>(v, in, tl, twN);
        } else {
            if (r == R)
                rsplit_combine<T, LOGM, LOGP, B, E, INV, R>(v, in, tl, twN);
            else
                RsplitDispatch<T, LOGM, LOGP, B, E, INV, R + 1, G>::go(r, v, in, tl, twN);
        }
    }
};

// Grid: 8 * G * ceil(ntrans / 8) work-groups (G = P / B per transform);
// work-group g -> XCD lane x = g mod 8, s = g / 8, transform
// t = (s / G) * 8 + x, residue group r = s mod G.
template <typename T, int LOGM, int LOGP, int B, int E, bool INV, int WPE>
__global__ void __launch_bounds__((1 << LOGM) / E, WPE) rsplit_kernel(TileArgs A) {
    using C2 = c2<T>;
    constexpr int M = 1 << LOGM, P = 1 << LOGP, G = P / B, TPT = M / E;
    constexpr unsigned N = (unsigned)M * (unsigned)P;
    static_assert(B == 1 || B == 2, "residues per work-group");
    static_assert(P >= 2 && P <= 16 && G >= 1, "split radix");
    extern __shared__ __align__(16) unsigned char smem_raw[];
    C2* slot = reinterpret_cast<C2*>(smem_raw);
    const C2* __restrict__ tw = reinterpret_cast<const C2*>(A.tw);
    const int tl = threadIdx.x;
    const unsigned g = blockIdx.x;
    const unsigned s = g >> 3;
    const unsigned t = (s / (unsigned)G) * 8u + (g & 7u);
    const int r = (int)(s % (unsigned)G);
    const TransformRef me = locate(A, t);
    if (!me.valid) return;  // uniform over the work-group
    const C2* __restrict__ in = reinterpret_cast<const C2*>(A.in) + me.ib;
    C2 v[B][E];
    RsplitDispatch<T, LOGM, LOGP, B, E, INV, 0, G>::go(r, v, in, tl, A.tw64 + N);
    Core<T, LOGM, E, INV>::run(v[0], slot, tl, tw);  // v[0][m]: X[P k + rB], k = tl + m*TPT
    if constexpr (B == 2) {
        __syncthreads();  // slot free for the second residue
        Core<T, LOGM, E, INV>::run(v[1], slot, tl, tw);
    }
    C2* __restrict__ out = reinterpret_cast<C2*>(A.out) + me.ob;
    const unsigned q0 = (unsigned)(r * B);
    if constexpr (B == 2 && sizeof(T) == 4) {
        // FP32 pairs (X[Pk + q0], X[Pk + q0 + 1]) as one 16-byte store where
        // the output address allows it (q0 and P are even)
        if ((reinterpret_cast<unsigned long long>(out) & 15ull) == 0) {
            float4* o4 = reinterpret_cast<float4*>(out);
#pragma unroll
            for (int m = 0; m < E; ++m) {
                const unsigned k = (unsigned)(tl + m * TPT);
                const unsigned i = (k * (unsigned)P + q0) >> 1;
                *reinterpret_cast<float4*>(reinterpret_cast<char*>(o4) + i * 16u) =
                    make_float4(v[0][m].x, v[0][m].y, v[1][m].x, v[1][m].y);
            }
            return;
        }
    }
#pragma unroll
    for (int m = 0; m < E; ++m) {
        const unsigned k = (unsigned)(tl + m * TPT);
#pragma unroll
        for (int b = 0; b < B; ++b) st_at<T>(out, k * (unsigned)P + q0 + (unsigned)b, v[b][m]);
    }
}

template <typename T, int LOGM>
constexpr size_t rsplit_lds_bytes() {
    return (size_t)slot_len(1 << LOGM) * sizeof(c2<T>);
}

}  // namespace mfx
