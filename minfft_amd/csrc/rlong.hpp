// rlong.hpp -- real, DCT and DST transforms whose half-length complex FFT
// (L = N/2 = N1*N2 points) is longer than one tile, in two passes over HBM:
//
//   step 1 (rstep1_kernel): the kind's pre-processing (pre_point, kinds.hpp)
//     computed while the input is read, the N1-point column FFTs over n1 of
//     z[n1*N2 + n2], and the two-step twiddle W_L^{n2*k1}; written to the
//     scratch as S[k1*N2 + n2];
//   step 2 (rstep2_kernel): the N2-point row FFTs over n2 of S[k1*N2 + .],
//     giving the FFT output at position q = k1 + N1*k2, and the kind's
//     post-processing (post_item, kinds.hpp, in FP64 from the FP64 twiddle
//     table, rounded once at the store) written straight to the output.
//
// The post-processing of the real DFT and of DCT-2/DST-2 pairs position q
// with L - q (minfft.c:547-578, 707-725, 756-774).  L - q lies in row
// N1 - k1 (row 0 and row N1/2 pair with themselves), so a step-2 tile holds
// row pairs: slots j < H = C/2 hold rows k1 = P0 + j, slots H + j their
// partners (N1 - k1, or N1/2 for k1 = 0).  The other kinds read one FFT
// output per output item; their tiles pair row k1 with N1 - 1 - k1, whose
// items interleave with row k1's in the output.  Each output item is
// written by the tile that holds its FFT position (the inverse maps below).
// Pre- and post-twiddles are FP64 products of two short table runs
// (TwFactor), rounded once.
//
// Step 1 reads the input once (DCT-2/DST-2 and DCT-4/DST-4 gather two
// reals per point, the inverse kinds two mirrored items: these reads stay
// in L2), step 2 writes every output once: 2 x (input + output) bytes
// instead of the 4 passes of pre-kernel, two-step FFT and post-kernel.
#pragma once
#include "tile.hpp"

namespace mfx {

// Twiddle W_M^m for table index M + m (M a power of two, m < M), as the FP64
// product W_{M/2^s}^{hi} * W_M^{lo} of two short runs of the FP64 table
// (m = hi*2^s + lo, s = ceil(log2 M / 2): a few tens of KiB that stay in
// the caches, instead of one table entry per output item), rounded once to
// T.  The product adds at most a few FP64 ulps.
template <typename T>
struct TwFactor {
    const double2* t;
    __device__ __forceinline__ c2<T> operator[](int idx) const {
        const int lgM = 31 - __builtin_clz((unsigned)idx);
        const int s = (lgM + 1) >> 1;
        const int m = idx - (1 << lgM);
        const double2 a = t[(1 << (lgM - s)) + (m >> s)];
        const double2 b = t[(1 << lgM) + (m & ((1 << s) - 1))];
        return mk<T>((T)(a.x * b.x - a.y * b.y), (T)(a.x * b.y + a.y * b.x));
    }
};

// Row / column pairing of the kinds: the real DFT and DCT-2/DST-2 read FFT
// outputs q and L - q together (row k1 with N1 - k1, N1/2 for k1 = 0); the
// others write neighbouring output items from positions q and L - 1 - q
// (row k1 with N1 - 1 - k1), and read neighbouring input items for points
// p and L - 1 - p or L - p.
template <int KIND>
constexpr bool rlong_mirror() {
    return KIND == K_REALDFT || KIND == K_DCT2 || KIND == K_DST2;
}

// Step 1: a tile is C adjacent columns n2 of one transform; lane t takes
// column t % C (neighbouring lanes read and write neighbouring points) and
// FFT thread t / C.  A.fs_logN = log2 L; the scratch (A.out) holds
// transform b at b*L.
template <typename T, int LOG1, int KIND>
__global__ void __launch_bounds__(512) rstep1_kernel(TileArgs A) {
    using C2 = c2<T>;
    constexpr int N1 = 1 << LOG1;
    constexpr int E = default_E(LOG1);
    constexpr int TPT = N1 / E;
    constexpr bool INV = kind_inverse_fft(KIND);
    extern __shared__ __align__(16) unsigned char smem_raw[];
    C2* lds = reinterpret_cast<C2*>(smem_raw);
    const C2* __restrict__ tw = reinterpret_cast<const C2*>(A.tw);

    const int t = threadIdx.x;
    const int logL = A.fs_logN;
    const int L = 1 << logL;
    const int N2 = L >> LOG1;
    const int lg_tiles = logL - LOG1 - A.lgC;  // log2 of the tiles per transform
    const unsigned tile = tile_id(A);
    const unsigned b = tile >> lg_tiles;
    const int c = t & (A.C - 1), tl = t >> A.lgC;
    // column blocks in mirror order (0, last, 1, last - 1, ...): the blocks
    // whose points read the same input lines run side by side
    const int u = (int)(tile & ((1u << lg_tiles) - 1u));
    const int blk = lg_tiles == 0 ? 0 : ((u & 1) ? (1 << lg_tiles) - 1 - (u >> 1) : (u >> 1));
    const int n2 = (blk << A.lgC) + c;
    const TransformRef r = locate(A, b);
    C2* slot = lds + (long long)c * A.sc;

    const T* __restrict__ xr = reinterpret_cast<const T*>(A.in);
    const C2* __restrict__ xc = reinterpret_cast<const C2*>(A.in);
    const long long es = A.in_e;
    auto X = [&](int i) -> T { return xr[r.ib + (long long)i * es]; };
    auto Z = [&](int i) -> C2 {
        if constexpr (KIND == K_REALDFT) {
            const long long q = r.ib + 2LL * i * es;
            return mk<T>(xr[q], xr[q + es]);
        } else {
            return xc[r.ib + (long long)i * es];
        }
    };
    const TwFactor<T> twf{A.tw64};
    C2 v[E];
    if (r.valid) {
#pragma unroll
        for (int m = 0; m < E; ++m) v[m] = pre_point<T, KIND>((tl + m * TPT) * N2 + n2, L, twf, X, Z);
    } else {
#pragma unroll
        for (int m = 0; m < E; ++m) v[m] = mk<T>(0, 0);
    }
    Core<T, LOG1, E, INV>::run(v, slot, tl, tw, false);

    // W_L^{n2*k1} on output k1 = tl + m*TPT, as W_L^{n2*tl} (W_L^{n2*TPT})^m in FP64
    const double2* twN = A.tw64 + L;
    double2 w = twN[((unsigned)n2 * (unsigned)tl) & (unsigned)(L - 1)];
    const double2 st = twN[((unsigned)n2 * (unsigned)TPT) & (unsigned)(L - 1)];
#pragma unroll
    for (int m = 0; m < E; ++m) {
        const C2 wt = mk<T>((T)w.x, (T)w.y);
        v[m] = INV ? cmulc<T>(v[m], wt) : cmul<T>(v[m], wt);
        if (m + 1 < E) w = make_double2(w.x * st.x - w.y * st.y, w.x * st.y + w.y * st.x);
    }
    if (r.valid) {
        C2* __restrict__ S = reinterpret_cast<C2*>(A.out) + (long long)b * L;
#pragma unroll
        for (int m = 0; m < E; ++m) S[(long long)(tl + m * TPT) * N2 + n2] = v[m];
    }
}

// Output items of the real DFT, DCT-2/DST-2 and the inverse real DFT whose
// post-processing is anchored at FFT position q (every item has exactly one
// anchor); returns their count.  (DCT-3/DST-3 and DCT-4/DST-4 take positions
// q and L - 1 - q together, in rstep2_kernel.)
template <int KIND>
__device__ __forceinline__ int rlong_items(int q, int L, int (&it)[2]) {
    const int N = 2 * L;
    if constexpr (KIND == K_REALDFT) {
        it[0] = q;  // X[0] and X[L] both come from position 0
        it[1] = L;
        return q == 0 ? 2 : 1;
    } else if constexpr (KIND == K_DCT2) {
        it[0] = q == 0 ? 0 : q;
        it[1] = q == 0 ? L : N - q;
        return 2;
    } else if constexpr (KIND == K_DST2) {
        it[0] = q == 0 ? N - 1 : q - 1;
        it[1] = q == 0 ? L - 1 : N - 1 - q;
        return 2;
    } else {  // K_INVREALDFT: item i reads position i >> 1
        static_assert(KIND == K_INVREALDFT, "anchored items of this kind");
        it[0] = 2 * q;
        it[1] = 2 * q + 1;
        return 2;
    }
}

// Step 2: a tile is C rows of one transform (row pairs, see above), with a
// row-mapped core (row t / TPT, thread t % TPT; a row's threads share a
// wavefront).  The rows' FFT outputs go to LDS; then lane t takes slot
// t % C and positions k2 = t / C + j*TPT, so that neighbouring lanes write
// neighbouring output items.  A.in is the scratch, A.out the
// output (through the batch levels / entry table of the pass).
template <typename T, int LOG2, int KIND>
__global__ void __launch_bounds__(512) rstep2_kernel(TileArgs A) {
    using C2 = c2<T>;
    constexpr int N2 = 1 << LOG2;
    constexpr int E = default_E(LOG2);
    constexpr int TPT = N2 / E;
    constexpr int LGT = ilog2(TPT);
    constexpr bool INV = kind_inverse_fft(KIND);
    extern __shared__ __align__(16) unsigned char smem_raw[];
    C2* lds = reinterpret_cast<C2*>(smem_raw);
    const C2* __restrict__ tw = reinterpret_cast<const C2*>(A.tw);

    const int t = threadIdx.x;
    const int logL = A.fs_logN;
    const int L = 1 << logL;
    const int LOG1 = logL - LOG2;
    const int N1 = 1 << LOG1;
    const int C = A.C, H = C >> 1;
    const int lg_tiles = LOG1 - A.lgC;  // N1 / C tiles per transform
    const unsigned tile = tile_id(A);
    const unsigned b = tile >> lg_tiles;
    const int P0 = (int)(tile & ((1u << lg_tiles) - 1u)) * H;
    auto row_of = [&](int s) -> int {
        const int k1 = P0 + (s & (H - 1));
        if (s < H) return k1;
        if constexpr (rlong_mirror<KIND>()) return k1 == 0 ? N1 / 2 : N1 - k1;
        return N1 - 1 - k1;
    };
    const TransformRef r = locate(A, b);

    // ---- row FFTs ----
    {
        const int c = t >> LGT, tl = t & (TPT - 1);
        C2* slot = lds + (long long)c * A.sc;
        const C2* __restrict__ S =
            reinterpret_cast<const C2*>(A.in) + (long long)(r.valid ? b : 0) * L + (long long)row_of(c) * N2;
        C2 v[E];
#pragma unroll
        for (int m = 0; m < E; ++m) v[m] = S[tl + m * TPT];
        Core<T, LOG2, E, INV>::run(v, slot, tl, tw, TPT <= 64);
        tile_sync(TPT <= 64);  // the row's last slot reads are done
        Core<T, LOG2, E, INV>::write_owned(v, slot, tl);
    }
    __syncthreads();
    if (!r.valid) return;

    // ---- post-processing, c-fast over the tile's rows ----
    const int s = t & (C - 1);
    const int row = row_of(s);
    const C2* my = lds + (long long)s * A.sc;
    const C2* partner = lds + (long long)(s ^ H) * A.sc;
    auto Tf = [&](int pos) -> double2 {
        const C2 z = ((pos & (N1 - 1)) == row ? my : partner)[lpad(pos >> LOG1)];
        return make_double2((double)z.x, (double)z.y);
    };
    const TwFactor<double> twf{A.tw64};
    const int per = blockDim.x >> A.lgC;  // positions k2 advanced per round
    T* __restrict__ yr = reinterpret_cast<T*>(A.out) + (kind_complex_out(KIND) ? 2 : 1) * r.ob;
    const long long oe = A.out_e;
    auto item = [&](int i) -> T { return (T)post_item<double, KIND>(i, L, twf, Tf).x; };
    if constexpr (KIND == K_DCT4 || KIND == K_DST4 || KIND == K_DCT3 || KIND == K_DST3) {
        // positions q < L/2 take their partner q' = L - 1 - q (the partner
        // row of the tile) along: DCT-4/DST-4 items 2q, 2q+1 and N-2-2q,
        // N-1-2q; DCT-3/DST-3 items 4q .. 4q+3 -- neighbouring items, stored
        // together when the output is contiguous and aligned
        constexpr bool T4 = (KIND == K_DCT4 || KIND == K_DST4);
        const bool vec = oe == 1 && (r.ob & (T4 ? 1 : 3)) == 0;
        const int N = 2 * L;
#pragma unroll 2
        for (int k2 = t >> A.lgC; k2 < N2; k2 += per) {
            const int q = row + (k2 << LOG1);
            if (q >= L - 1 - q) continue;
            if constexpr (T4) {
                const T a0 = item(2 * q), a1 = item(2 * q + 1);
                const T b0 = item(N - 2 - 2 * q), b1 = item(N - 1 - 2 * q);
                if (vec) {
                    *reinterpret_cast<c2<T>*>(yr + 2 * q) = mk<T>(a0, a1);
                    *reinterpret_cast<c2<T>*>(yr + (N - 2 - 2 * q)) = mk<T>(b0, b1);
                } else {
                    yr[(long long)(2 * q) * oe] = a0;
                    yr[(long long)(2 * q + 1) * oe] = a1;
                    yr[(long long)(N - 2 - 2 * q) * oe] = b0;
                    yr[(long long)(N - 1 - 2 * q) * oe] = b1;
                }
            } else {
                T v[4];
#pragma unroll
                for (int h = 0; h < 4; ++h) v[h] = item(4 * q + h);
                if (vec) {
                    if constexpr (sizeof(T) == 4) {
                        *reinterpret_cast<float4*>(yr + 4 * q) = make_float4(v[0], v[1], v[2], v[3]);
                    } else {
                        *reinterpret_cast<double2*>(yr + 4 * q) = make_double2(v[0], v[1]);
                        *reinterpret_cast<double2*>(yr + 4 * q + 2) = make_double2(v[2], v[3]);
                    }
                } else {
#pragma unroll
                    for (int h = 0; h < 4; ++h) yr[(long long)(4 * q + h) * oe] = v[h];
                }
            }
        }
    } else {
#pragma unroll 4
        for (int k2 = t >> A.lgC; k2 < N2; k2 += per) {
            const int q = row + (k2 << LOG1);
            int items[2];
            const int n = rlong_items<KIND>(q, L, items);
            if constexpr (KIND == K_DCT2 || KIND == K_DST2) {
                // both items from one real-DFT value and one twiddle
                double a, bb;
                t2_post_pair<double, KIND == K_DST2>(q, L, twf, Tf, a, bb);
                yr[(long long)items[0] * oe] = (T)a;
                yr[(long long)items[1] * oe] = (T)bb;
            } else {
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    if (h >= n) break;
                    const double2 yd = post_item<double, KIND>(items[h], L, twf, Tf);
                    if constexpr (kind_complex_out(KIND)) {
                        reinterpret_cast<C2*>(yr)[(long long)items[h] * oe] = mk<T>((T)yd.x, (T)yd.y);
                    } else {
                        yr[(long long)items[h] * oe] = (T)yd.x;
                    }
                }
            }
        }
    }
}

// log2 N1 / log2 N2 of the instantiated step kernels
constexpr int kRlongLogMin = 7;
constexpr int kRlongLogMax = 12;

}  // namespace mfx
