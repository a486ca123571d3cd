// tile.hpp -- the fused tile kernel: C independent 1D transforms of length
// L = 2^LOGL per work-group, with the kind's pre-processing fused in front of
// the FFT core and its post-processing fused behind it.
//
// Addressing.  A pass covers a batch of transforms described by up to four
// batch levels (level 3 varies fastest).  Transform b of the batch has index
// (i0,i1,i2,i3) and its items sit at
//     base + e * estride,   base = sum_l i_l * stride_l   (or, for level 0,
//     an entry-offset table when one is given: batched uneven workloads),
// separately for the input and the output side, in units of that side's item
// (complex point for complex sides, real value for real sides).
//
// Lane mapping.  Either "row" mapping (the TPT threads of one transform are
// adjacent lanes: coalesced when estride == 1) or "c-fast" mapping (lane
// neighbours work on neighbouring transforms: coalesced when the fastest
// batch level has stride 1, i.e. strided columns of a multi-dimensional
// array).  Each side is either read/written directly from registers with
// the core's mapping, or staged through LDS with its own mapping.
#pragma once
#include "kinds.hpp"

namespace mfx {

enum IoMode : int { IO_DIRECT = 0, IO_STAGE_ROW = 1, IO_STAGE_CF = 2 };

struct TileArgs {
    const void* in;
    void* out;
    const void* tw;                    // twiddle table in the kernel's precision
    const double2* tw64;               // FP64 master table (two-step twiddles)
    const long long* in_off;           // optional level-0 entry offsets (input side)
    const long long* out_off;          // optional level-0 entry offsets (output side)
    long long in_s[4], out_s[4];       // batch strides
    long long in_e, out_e;             // item strides
    unsigned B[4];                     // batch extents
    int lgB[4];                        // log2 B[l] when B[l] is a power of two, else -1
    unsigned ntrans;                   // product of B
    unsigned ntiles;                   // grid size
    int C;                             // transforms per tile (a power of two)
    int lgC;                           // log2 C
    int cfast;                         // core lane mapping: 1 = c-fast
    int in_mode, out_mode;             // IoMode
    int fs_level;                      // two-step twiddle: batch level holding n2
    int fs_logN;                       // 0: none; else multiply by W_{2^fs_logN}^{n2*k}
    int sc;                            // LDS slot length (complex points)
    int xcd;                           // number of XCDs for the tile swizzle (1 = off)
    int wave;                          // 1: row-mapped core, so a transform's lanes
                                       // share one wavefront when TPT <= 64
    long long pack_in, pack_out;       // packed-pair tiles: item offset of column
                                       // H-1 from column 0 (tile_kernel PACK)
    unsigned* hflag;                   // one-work-group launches: host-polled completion
    unsigned hval;                     // word (pinned memory) and the value to store there
};

struct TransformRef {
    long long ib, ob;
    unsigned i0, i1, i2, i3;
    bool valid;
    // batch index at level l (a select chain: no private-memory array)
    __device__ __forceinline__ unsigned at(int l) const {
        return l == 3 ? i3 : (l == 2 ? i2 : (l == 1 ? i1 : i0));
    }
};

// n / B[l] and n % B[l]: shifts for power-of-two extents (the common case;
// the branch is uniform), a division otherwise
__device__ __forceinline__ unsigned divmod_level(const TileArgs& A, int l, unsigned n, unsigned& rem) {
    if (A.lgB[l] >= 0) {
        rem = n & (A.B[l] - 1u);
        return n >> A.lgB[l];
    }
    const unsigned q = n / A.B[l];
    rem = n - q * A.B[l];
    return q;
}

__device__ __forceinline__ TransformRef locate(const TileArgs& A, unsigned b) {
    TransformRef r;
    r.valid = b < A.ntrans;
    if (!r.valid) b = 0;
    unsigned i3, i2, i1;
    unsigned t = divmod_level(A, 3, b, i3);
    t = divmod_level(A, 2, t, i2);
    const unsigned i0 = divmod_level(A, 1, t, i1);
    r.i0 = i0; r.i1 = i1; r.i2 = i2; r.i3 = i3;
    long long ib = (A.in_off ? A.in_off[i0] : (long long)i0 * A.in_s[0]) +
                   (long long)i1 * A.in_s[1] + (long long)i2 * A.in_s[2] + (long long)i3 * A.in_s[3];
    long long ob = (A.out_off ? A.out_off[i0] : (long long)i0 * A.out_s[0]) +
                   (long long)i1 * A.out_s[1] + (long long)i2 * A.out_s[2] + (long long)i3 * A.out_s[3];
    r.ib = ib;
    r.ob = ob;
    return r;
}

// Remap the hardware work-group id so that consecutive tiles (neighbouring
// columns in memory) land on the same XCD and share its L2: work-groups are
// dispatched round-robin over the XCDs.
__device__ __forceinline__ unsigned tile_id(const TileArgs& A) {
    unsigned g = blockIdx.x;
    const unsigned X = (unsigned)A.xcd;
    if (X <= 1) return g;
    const unsigned per = A.ntiles / X;
    if (g >= per * X) return g;  // tail keeps identity
    return (g % X) * per + g / X;
}

// Items a + s*(stl + k*TPT) (s = +-1) of a strided row, k = 0, 1, ...: one
// 64-bit base per family of positions and a uniform step, instead of a
// 64-bit multiply per item.
template <typename P>
struct Walk {
    P* p0;
    long long step;
    __device__ __forceinline__ Walk(P* base, int a, int s, int stl, int TPT, long long es)
        : p0(base + (long long)(a + s * stl) * es), step((long long)s * TPT * es) {}
    __device__ __forceinline__ P& operator[](int k) const { return p0[k * step]; }
};

// 16-byte vectors of reals, for contiguous real rows
template <typename T> struct Vec16;
template <> struct Vec16<float> { using V = float4; static constexpr int n = 4; };
template <> struct Vec16<double> { using V = double2; static constexpr int n = 2; };

// real item i of a slot that holds real items: the pair (2p, 2p+1) sits at
// the padded complex position of p
__device__ __forceinline__ int ridx(int i) { return 2 * lpad(i >> 1) + (i & 1); }

// Staged input: this lane's share (items stl, stl + TPT, ...) of transform
// `src` into its LDS slot.  Trip counts are compile-time constants, so the
// loads of a lane are all issued before the first LDS store waits on them;
// contiguous 16-byte-aligned real rows use 16-byte loads.
template <typename T, int KIND, int L, int TPT>
__device__ __forceinline__ void stage_in(const TileArgs& A, const TransformRef& src, c2<T>* sslot, int stl) {
    using C2 = c2<T>;
    if constexpr (kind_complex_in(KIND)) {
        const C2* __restrict__ in = reinterpret_cast<const C2*>(A.in) + src.ib;
        constexpr int K = L / TPT;
        const long long es = A.in_e;
        const Walk<const C2> wi(in, 0, 1, stl, TPT, es);
        C2 r[K];
#pragma unroll
        for (int k = 0; k < K; ++k) r[k] = wi[k];
        C2* d0 = sslot + lpad(stl);
#pragma unroll
        for (int k = 0; k < K; ++k) {
            if constexpr (TPT % 16 == 0) d0[k * (TPT + TPT / 16)] = r[k];
            else sslot[lpad(stl + k * TPT)] = r[k];
        }
        if constexpr (kind_nin(KIND, L) > L) {  // inverse real DFT: item L
            if (stl == 0) sslot[lpad(L)] = in[(long long)L * es];
        }
    } else {
        // 2L real items (the real DFT's input pairs, or a DCT/DST row)
        constexpr int NR = 2 * L;
        using V = typename Vec16<T>::V;
        constexpr int VN = Vec16<T>::n;
        const T* __restrict__ in = reinterpret_cast<const T*>(A.in) + src.ib;
        T* rs = reinterpret_cast<T*>(sslot);
        if constexpr (NR % (VN * TPT) == 0) {
            if (A.in_e == 1 && (reinterpret_cast<unsigned long long>(in) & 15) == 0) {
                constexpr int K = NR / (VN * TPT);
                V r[K];
#pragma unroll
                for (int k = 0; k < K; ++k) r[k] = reinterpret_cast<const V*>(in)[stl + k * TPT];
                // first complex position p = (VN/2)(stl + k TPT); padded
                // position linear in k when (VN/2) TPT is a multiple of 16
                constexpr int PS = (VN / 2) * TPT;
                C2* d0 = sslot + lpad((VN / 2) * stl);
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    C2* d = (PS % 16 == 0) ? d0 + k * (PS + PS / 16) : sslot + lpad((VN / 2) * (stl + k * TPT));
                    if constexpr (VN == 4) {
                        // positions p, p+1 (p even: never split by a pad gap)
                        d[0] = mk<T>(r[k].x, r[k].y);
                        d[1] = mk<T>(r[k].z, r[k].w);
                    } else {
                        d[0] = mk<T>(r[k].x, r[k].y);
                    }
                }
                return;
            }
        }
        constexpr int K = NR / TPT;
        const Walk<const T> wi(in, 0, 1, stl, TPT, A.in_e);
        T r[K];
#pragma unroll
        for (int k = 0; k < K; ++k) r[k] = wi[k];
        // real item i = stl + k TPT sits at ridx(i); linear in k when TPT/2
        // is a multiple of 16 (TPT even keeps the parity of i)
        T* d0 = rs + ridx(stl);
#pragma unroll
        for (int k = 0; k < K; ++k) {
            if constexpr (TPT % 32 == 0) d0[k * (TPT + TPT / 16)] = r[k];
            else rs[ridx(stl + k * TPT)] = r[k];
        }
    }
}

// Staged output: the kind's output items of transform `dst` from the FFT
// output in its LDS slot (post-processing fused), this lane's share.
template <typename T, int KIND, int L, int TPT>
__device__ __forceinline__ void stage_out(const TileArgs& A, const TransformRef& dst, const c2<T>* sslot, int stl,
                                          const c2<T>* __restrict__ tw) {
    using C2 = c2<T>;
    auto Tf = [&](int p) -> C2 { return sslot[lpad(p)]; };
    if constexpr (kind_complex_out(KIND)) {
        C2* __restrict__ out = reinterpret_cast<C2*>(A.out) + dst.ob;
        constexpr int K = L / TPT;
        const long long es = A.out_e;
        const Walk<C2> wo(out, 0, 1, stl, TPT, es);
#pragma unroll
        for (int k = 0; k < K; ++k) wo[k] = post_item<T, KIND>(stl + k * TPT, L, tw, Tf);
        if constexpr (kind_nout(KIND, L) > L) {  // real DFT: item L
            if (stl == 0) out[(long long)L * es] = post_item<T, KIND>(L, L, tw, Tf);
        }
    } else {
        constexpr int NR = 2 * L;
        using V = typename Vec16<T>::V;
        constexpr int VN = Vec16<T>::n;
        T* __restrict__ out = reinterpret_cast<T*>(A.out) + dst.ob;
        if constexpr (NR % (VN * TPT) == 0) {
            if (A.out_e == 1 && (reinterpret_cast<unsigned long long>(out) & 15) == 0) {
                constexpr int K = NR / (VN * TPT);
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    const int j = stl + k * TPT;
                    T w[VN];
#pragma unroll
                    for (int q = 0; q < VN; ++q) w[q] = post_item<T, KIND>(VN * j + q, L, tw, Tf).x;
                    V o;
                    if constexpr (VN == 4) o = make_float4(w[0], w[1], w[2], w[3]);
                    else o = make_double2(w[0], w[1]);
                    reinterpret_cast<V*>(out)[j] = o;
                }
                return;
            }
        }
        constexpr int K = NR / TPT;
        const Walk<T> wo(out, 0, 1, stl, TPT, A.out_e);
#pragma unroll
        for (int k = 0; k < K; ++k) wo[k] = post_item<T, KIND>(stl + k * TPT, L, tw, Tf).x;
    }
}

// ---------------------------------------------------------------------------
// Pre- and post-processing of the real kinds inside the tile kernel, in
// symmetric groups so that the unrolled loops carry no per-item special
// cases: the only branches are for the lane holding point 0 (and L/2).  The
// formulas are those of kinds.hpp (same references), regrouped:
//   * the C2R pre-twiddle (inverse real DFT, DCT-3, DST-3) of points p and
//     L-p shares z[m], z[L-m] and W_N^m, m = min(p, L-p);
//   * the R2C post-twiddle (real DFT, DCT-2, DST-2) of group m gives Z[m]
//     and Z[L-m] from T[m], T[L-m] and W_N^m; DCT-2/DST-2 turn each Z[n]
//     into two outputs (real and imaginary part of Z[n]*W_{4N}^n);
//   * DCT-4/DST-4 and the un-reorders produce output pairs (2j, 2j+1).
// Used for E >= 2 (L >= 2); length-1 FFTs use the per-item formulas.
// ---------------------------------------------------------------------------

// FFT input points tl + m*TPT of this lane, from the kind's input staged in
// its own LDS slot (real items at ridx, complex items at lpad)
template <typename T, int KIND, int L, int E, int TPT>
__device__ __forceinline__ void pre_owned(c2<T>* v, const c2<T>* slot, int tl, const c2<T>* __restrict__ tw) {
    using C2 = c2<T>;
    static_assert(E >= 2 && E % 2 == 0, "pre_owned needs E >= 2");
    constexpr int N = 2 * L, H = L / 2;
    const T* rs = reinterpret_cast<const T*>(slot);
    auto X = [&](int i) -> T { return rs[ridx(i)]; };
    if constexpr (KIND == K_DCT2 || KIND == K_DST2) {
        // t[n] = x[2n], t[L+n] = s*x[N-1-2n]; point p = (t[2p], t[2p+1])
        constexpr T sg = (KIND == K_DST2) ? T(-1) : T(1);
        if constexpr (TPT % 8 == 0) {
            // x[4p], x[4p+2] are the .x of complex positions 2p, 2p+1;
            // x[N-1-4q], x[N-3-4q] the .y of L-1-2q, L-2-2q: padded
            // positions linear in m (2*TPT is a multiple of 16)
            constexpr int S = 2 * TPT + (2 * TPT) / 16;
            const C2* lo = slot + lpad(2 * tl);
            const C2* hi = slot + lpad(L - 1 - 2 * tl);
#pragma unroll
            for (int m = 0; m < E; ++m) {
                if (m < E / 2) {
                    v[m] = mk<T>(lo[m * S].x, lo[m * S + 1].x);
                } else {
                    const int mq = m - E / 2;
                    v[m] = mk<T>(sg * hi[-mq * S].y, sg * hi[-mq * S - 1].y);
                }
            }
        } else {
#pragma unroll
            for (int m = 0; m < E; ++m) {
                const int p = tl + m * TPT;
                if (m < E / 2) {
                    v[m] = mk<T>(X(4 * p), X(4 * p + 2));
                } else {
                    const int q = p - H;
                    v[m] = mk<T>(sg * X(N - 1 - 4 * q), sg * X(N - 3 - 4 * q));
                }
            }
        }
    } else if constexpr (KIND == K_DCT4 || KIND == K_DST4) {
        // x[2p] is the .x of complex position p, x[N-1-2p] the .y of L-1-p
        constexpr int S = TPT + TPT / 16;
        const C2* lo = slot + lpad(tl);
        const C2* hi = slot + lpad(L - 1 - tl);
#pragma unroll
        for (int m = 0; m < E; ++m) {
            const int p = tl + m * TPT;
            const C2 e = twl<T>(tw, (unsigned)(2 * N + p));
            T xa, xb;
            if constexpr (TPT % 16 == 0) {
                xa = lo[m * S].x;
                xb = hi[-m * S].y;
            } else {
                xa = X(2 * p);
                xb = X(N - 1 - 2 * p);
            }
            if constexpr (KIND == K_DCT4)
                v[m] = mk<T>(e.x * xa - e.y * xb, e.x * xb + e.y * xa);
            else
                v[m] = mk<T>(-e.x * xa - e.y * xb, e.x * xb - e.y * xa);
        }
    } else {  // inverse real DFT, DCT-3, DST-3: C2R pre-twiddle of z
        // z[0..L] sits in the slot (the inverse real DFT's input; for
        // DCT-3/DST-3 put there by t3_to_z)
        auto zg = [&](int k) -> C2 { return slot[lpad(k)]; };
        constexpr int S = TPT + TPT / 16;
        const C2* zp = slot + lpad(tl);      // z[p] at m*S (TPT a multiple of 16)
        const C2* zq = slot + lpad(L - tl);  // z[L-p] at -m*S
#pragma unroll
        for (int m = 0; m < E; ++m) {
            const int p = tl + m * TPT;
            const bool lower = m < E / 2;  // p < L/2
            // mm = min(p, L-p) in [0, L/2]: 0 only for point 0, whose result
            // is replaced below (its reads, z[0] and z[L], are in the slot)
            const int mm = lower ? p : L - p;
            C2 a, b;
            if constexpr (TPT % 16 == 0) {
                a = lower ? zp[m * S] : zq[-m * S];
                b = lower ? zq[-m * S] : zp[m * S];
            } else {
                a = zg(mm);
                b = zg(L - mm);
            }
            const C2 e = twl<T>(tw, (unsigned)(2 * L + mm));
            const T ur = a.x + b.x, ui = a.y - b.y;
            const T ttr = a.x - b.x, tti = a.y + b.y;
            const T vr = ttr * e.y - tti * e.x;  // (ttr + i tti) * i conj(e)
            const T vi = ttr * e.x + tti * e.y;
            C2 r = lower ? mk<T>(ur + vr, ui + vi) : mk<T>(ur - vr, -ui + vi);
            if (m == 0) {
                // point 0 from z[0], z[L] (imaginary parts ignored), selected
                // by the lane that holds it
                const T a0 = zg(0).x, b0 = zg(L).x;
                r = (tl == 0) ? mk<T>(a0 + b0, a0 - b0) : r;
            }
            if (m == E / 2) {
                // point L/2: 2 conj(z[L/2])
                const C2 zh = zg(H);
                r = (tl == 0) ? mk<T>(2 * zh.x, -2 * zh.y) : r;
            }
            v[m] = r;
        }
    }
}

// DCT-3 / DST-3: the staged real input x (2L items) -> z[0..L] in the same
// slot (kinds.hpp t3_z, minfft.c:807-817, 857-867).  Group g < L/2 of lane tl
// (g = tl + k*TPT) reads x at g, N-g, L-g, L+g (or their DST-3 mirrors) and
// yields z[g], z[L-g]; the lane of g = 0 yields z[0], z[L], z[L/2].  All
// reads happen before the barrier, all writes after it.
template <typename T, int KIND, int L, int E, int TPT>
__device__ __forceinline__ void t3_to_z(c2<T>* slot, int tl, const c2<T>* __restrict__ tw, bool wave) {
    using C2 = c2<T>;
    static_assert(E >= 2 && E % 2 == 0, "t3_to_z needs E >= 2");
    constexpr int N = 2 * L, H = L / 2;
    constexpr bool DS = (KIND == K_DST3);
    const T* rs = reinterpret_cast<const T*>(slot);
    auto X = [&](int i) -> T { return rs[ridx(i)]; };
    // z[k], 0 < k < L
    auto zg = [&](int k) -> C2 {
        const C2 e = twl<T>(tw, (unsigned)(4 * N + k));
        const T xa = DS ? X(N - k - 1) : X(k);
        const T xb = DS ? X(k - 1) : X(N - k);
        return mk<T>(xa * e.x - xb * e.y, -xa * e.y - xb * e.x);
    };
    constexpr int G = E / 2;
    C2 za[G], zb[G];
#pragma unroll
    for (int k = 0; k < G; ++k) {
        const int g = max(tl + k * TPT, 1);  // g = 0 is replaced below
        za[k] = zg(g);
        zb[k] = zg(L - g);
    }
    C2 zh = mk<T>(0, 0);
    const bool first = (tl == 0);
    if (first) {
        za[0] = mk<T>(DS ? X(N - 1) : X(0), (T)0);
        zb[0] = mk<T>(sqrt2_v<T>() * (DS ? X(L - 1) : X(L)), (T)0);
        zh = zg(H);
    }
    tile_sync(wave);
#pragma unroll
    for (int k = 0; k < G; ++k) {
        const int g = tl + k * TPT;
        slot[lpad(g)] = za[k];
        slot[lpad(L - g)] = zb[k];
    }
    if (first) slot[lpad(H)] = zh;
    tile_sync(wave);
}

// two real outputs (2j, 2j+1) of a row, p = &item 2j; one 2-wide store when
// contiguous
template <typename T>
__device__ __forceinline__ void store_pair(T* __restrict__ p, long long es, bool vec, T a, T b) {
    if (vec) {
        *reinterpret_cast<c2<T>*>(p) = mk<T>(a, b);
    } else {
        p[0] = a;
        p[es] = b;
    }
}

// the kind's outputs of transform dst from the FFT output in its LDS slot
template <typename T, int KIND, int L, int E, int TPT>
__device__ __forceinline__ void post_real(const TileArgs& A, const TransformRef& dst, const c2<T>* sslot, int stl,
                                          const c2<T>* __restrict__ tw) {
    using C2 = c2<T>;
    static_assert(E >= 2 && E % 2 == 0, "post_real needs E >= 2");
    constexpr int N = 2 * L, H = L / 2;
    auto Tf = [&](int p) -> C2 { return sslot[lpad(p)]; };
    const long long es = A.out_e;
    if constexpr (KIND == K_REALDFT || KIND == K_DCT2 || KIND == K_DST2) {
        // Z[n] -> outputs.  Real DFT: out[n] = Z[n].  DCT-2 / DST-2: with
        // c = W_{4N}^n, Re(Z c) goes to item re(n), -Im(Z c) to item im(n).
        constexpr bool DS = (KIND == K_DST2);
        auto re_pos = [&](int n) -> int { return DS ? N - 1 - n : n; };
        auto im_pos = [&](int n) -> int { return DS ? n - 1 : N - n; };
        C2* __restrict__ oc = reinterpret_cast<C2*>(A.out) + dst.ob;
        T* __restrict__ orl = reinterpret_cast<T*>(A.out) + dst.ob;
        // output items of group m as walks over k (m = stl + k*TPT):
        //   real DFT: Z[m] at m, Z[L-m] at L-m;
        //   DCT-2: re(m) = m, im(m) = N-m, re(L-m) = L-m, im(L-m) = L+m;
        //   DST-2: re(m) = N-1-m, im(m) = m-1, re(L-m) = L-1+m, im(L-m) = L-1-m.
        const Walk<C2> zm(oc, 0, 1, stl, TPT, es), zl(oc, L, -1, stl, TPT, es);
        const Walk<T> rm(orl, DS ? N - 1 : 0, DS ? -1 : 1, stl, TPT, es);
        const Walk<T> im_(orl, DS ? -1 : N, DS ? 1 : -1, stl, TPT, es);
        const Walk<T> rl(orl, DS ? L - 1 : L, DS ? 1 : -1, stl, TPT, es);
        const Walk<T> il(orl, DS ? L - 1 : L, DS ? -1 : 1, stl, TPT, es);
        // T[m] at tm[k*S], T[L-m] at tq[-k*S] (TPT a multiple of 16)
        constexpr int S = TPT + TPT / 16;
        const C2* tm = sslot + lpad(stl);
        const C2* tq = sslot + lpad(L - stl);
        // Re(Z c) and -Im(Z c) with c = W_{4N}^n
        auto rot = [&](C2 z, int n, T& re, T& im) {
            const C2 c = twl<T>(tw, (unsigned)(4 * N + n));
            re = 2 * (z.x * c.x - z.y * c.y);
            im = -2 * (z.x * c.y + z.y * c.x);
        };
#pragma unroll
        for (int k = 0; k < E / 2; ++k) {
            const int m = stl + k * TPT;  // group (m, L-m), 0 <= m < L/2
            if (k == 0 && stl == 0) {
                const C2 u = Tf(0);
                const C2 zh = conj<T>(Tf(H));
                if constexpr (KIND == K_REALDFT) {
                    oc[0] = mk<T>(u.x + u.y, (T)0);
                    oc[(long long)L * es] = mk<T>(u.x - u.y, (T)0);
                    oc[(long long)H * es] = zh;
                } else {
                    orl[(long long)re_pos(0) * es] = 2 * (u.x + u.y);
                    orl[(long long)re_pos(L) * es] = sqrt2_v<T>() * (u.x - u.y);
                    T re, im;
                    rot(zh, H, re, im);
                    orl[(long long)re_pos(H) * es] = re;
                    orl[(long long)im_pos(H) * es] = im;
                }
                continue;
            }
            C2 a, b;
            if constexpr (TPT % 16 == 0) {
                a = tm[k * S];
                b = tq[-k * S];
            } else {
                a = Tf(m);
                b = Tf(L - m);
            }
            const C2 e = twl<T>(tw, (unsigned)(2 * L + m));
            const T ur = (a.x + b.x) / 2, ui = (a.y - b.y) / 2;
            const T ttr = a.x - b.x, tti = a.y + b.y;
            const T vr = (ttr * e.y + tti * e.x) / 2;  // (ttr + i tti) * e / i / 2
            const T vi = (-ttr * e.x + tti * e.y) / 2;
            const C2 z0 = mk<T>(ur + vr, ui + vi), z1 = mk<T>(ur - vr, -ui + vi);  // Z[m], Z[L-m]
            if constexpr (KIND == K_REALDFT) {
                zm[k] = z0;
                zl[k] = z1;
            } else {
                T re, im;
                rot(z0, m, re, im);
                rm[k] = re;
                im_[k] = im;
                rot(z1, L - m, re, im);
                rl[k] = re;
                il[k] = im;
            }
        }
    } else {
        T* __restrict__ out = reinterpret_cast<T*>(A.out) + dst.ob;
        const bool vec = es == 1 && (reinterpret_cast<unsigned long long>(out) & (2 * sizeof(T) - 1)) == 0;
        const Walk<T> w2(out, 0, 1, stl, TPT, 2 * es);  // item 2j
        // linear padded positions (TPT a multiple of 32): T[j] at pj_[k*S],
        // T[L-1-j] at pq_[-k*S]; T[j>>1] at ph_[k*S2], T[(N-1-j)>>1] at pr_[-k*S2]
        constexpr int S = TPT + TPT / 16, S2 = TPT / 2 + TPT / 32;
        const C2* pj_ = sslot + lpad(stl);
        const C2* pq_ = sslot + lpad(L - 1 - stl);
        const C2* ph_ = sslot + lpad(stl >> 1);
        const C2* pr_ = sslot + lpad((N - 1 - stl) >> 1);
        constexpr bool LIN = (TPT % 32 == 0);
#pragma unroll
        for (int k = 0; k < E; ++k) {
            const int j = stl + k * TPT;  // 0 <= j < L
            T* pj = &w2[k];
            if constexpr (KIND == K_INVREALDFT) {
                const C2 t = LIN ? pj_[k * S] : Tf(j);
                store_pair<T>(pj, es, vec, t.x, t.y);
            } else if constexpr (KIND == K_DCT3 || KIND == K_DST3) {
                // y[2j] = u[j], y[2j+1] = s*u[N-1-j], u = FFT output as reals
                // (j and N-1-j keep the parity of stl and N-1-stl: TPT is even)
                constexpr T sg = (KIND == K_DST3) ? T(-1) : T(1);
                const int r = N - 1 - j;
                const C2 t0 = LIN ? ph_[k * S2] : Tf(j >> 1);
                const C2 t1 = LIN ? pr_[-k * S2] : Tf(r >> 1);
                store_pair<T>(pj, es, vec, (j & 1) ? t0.y : t0.x, sg * ((r & 1) ? t1.y : t1.x));
            } else {  // DCT-4 / DST-4
                const C2 e0 = twl<T>(tw, (unsigned)(8 * N + 4 * j + 1)), e1 = twl<T>(tw, (unsigned)(8 * N + 4 * j + 3));
                const C2 t0 = LIN ? pj_[k * S] : Tf(j);
                const C2 t1 = LIN ? pq_[-k * S] : Tf(L - 1 - j);
                T y0, y1;
                if constexpr (KIND == K_DCT4) {
                    y0 = 2 * (e0.x * t0.x - e0.y * t0.y);
                    y1 = 2 * (e1.x * t1.x + e1.y * t1.y);
                } else {
                    y0 = 2 * (e0.x * t0.y + e0.y * t0.x);
                    y1 = 2 * (-e1.x * t1.y + e1.y * t1.x);
                }
                store_pair<T>(pj, es, vec, y0, y1);
            }
        }
    }
}

// Largest work-group of a tile kernel: FP64 radix-16 butterflies need more
// than the 128 VGPRs a 1024-thread group leaves each lane.
template <typename T> constexpr int max_threads() { return sizeof(T) == 8 ? 512 : 1024; }
// radix-32 tiles (E = 32) hold twice the points per lane: at most 512 lanes
template <typename T, int E> constexpr int tile_threads() { return E >= 32 ? 512 : max_threads<T>(); }

// PACK (complex pass over a real DFT's half spectrum, direct c-fast I/O on
// both sides): see `packed` below.  NT: the direct complex I/O uses
// non-temporal accesses (passes over arrays larger than the MALL, which
// stream through it anyway).
template <typename T, int LOGL, int E, int KIND, bool PACK = false, bool NT = false>
__device__ __forceinline__ void tile_body(const TileArgs& A) {
    using C2 = c2<T>;
    constexpr int L = 1 << LOGL;
    constexpr int TPT = L / E;
    constexpr bool INV = kind_inverse_fft(KIND);
    constexpr bool CIN = kind_complex_in(KIND);
    constexpr bool COUT = kind_complex_out(KIND);
    constexpr int NIN = kind_nin(KIND, L);
    constexpr int NOUT = kind_nout(KIND, L);
    extern __shared__ __align__(16) unsigned char smem_raw[];
    C2* lds = reinterpret_cast<C2*>(smem_raw);
    const C2* __restrict__ tw = reinterpret_cast<const C2*>(A.tw);

    const int t = threadIdx.x;
    const int C = A.C;
    const unsigned tile = tile_id(A);
    // the two lane mappings
    const int c_row = t / TPT, tl_row = t - c_row * TPT;
    const int c_cf = t & (C - 1), tl_cf = t >> A.lgC;
    const int c = A.cfast ? c_cf : c_row;
    const int tl = A.cfast ? tl_cf : tl_row;
    const TransformRef me = locate(A, tile * (unsigned)C + (unsigned)c);
    C2* slot = lds + (long long)c * A.sc;
    // the core's exchanges stay inside a wavefront: no work-group barriers;
    // staging with the column mapping still needs them
    const bool wv = (TPT <= 64) && A.wave;
    const bool wv_in = wv && A.in_mode != IO_STAGE_CF;
    const bool wv_out = wv && A.out_mode != IO_STAGE_CF;

    C2 v[E];
    // Packed pair: the transform at h = 0 also carries column h = H-1 of a
    // real DFT's half spectrum, and the pass skips h = H-1.
    // Forward (after the real DFT along the last axis): both columns are real
    // (the DC and Nyquist terms of every row), so z = a + i b and, with
    // Z = DFT(z), A[k] = (Z[k] + conj Z[-k]) / 2, B[k] = (Z[k] - conj Z[-k]) / 2i.
    // Inverse (the pass whose output the inverse real DFT along the last axis
    // reads, which keeps only the real parts of h = 0 and h = H-1):
    // Re IDFT(c) = IDFT(c_h) with the Hermitian part c_h[k] = (c[k] + conj c[-k]) / 2,
    // and IDFT(a_h + i b_h) = Re IDFT(a) + i Re IDFT(b).
    static_assert(!PACK || KIND == K_DFT || KIND == K_INVDFT, "packed pairs: complex passes");
    const bool packed = PACK && me.valid && me.i3 == 0u;

    // ---------------- input side ----------------
    // (direct input only for the complex kinds: the real kinds always stage,
    // and compiling a dead direct path costs them registers)
    constexpr bool CPLX = (KIND == K_DFT || KIND == K_INVDFT);
    if (CPLX && A.in_mode == IO_DIRECT) {
        // complex input read straight into the owned positions
        const C2* __restrict__ in = reinterpret_cast<const C2*>(A.in) + me.ib;
        if (me.valid) {
            if (KIND == K_DFT && packed) {
                // real parts of the two columns (reals at twice the item stride)
                const T* __restrict__ ra = reinterpret_cast<const T*>(in + (long long)tl * A.in_e);
                const T* __restrict__ rb = reinterpret_cast<const T*>(in + A.pack_in + (long long)tl * A.in_e);
                const long long step = 2 * A.in_e * TPT;
#pragma unroll
                for (int m = 0; m < E; ++m) v[m] = mk<T>(ra[m * step], rb[m * step]);
            } else if (A.in_e == 1) {
#pragma unroll
                for (int m = 0; m < E; ++m) v[m] = NT ? ld_nt<T>(&in[tl + m * TPT]) : in[tl + m * TPT];
            } else {
                const long long es = A.in_e;
                const C2* __restrict__ b = in + (long long)tl * es;
                const long long step = es * TPT;
#pragma unroll
                for (int m = 0; m < E; ++m) v[m] = NT ? ld_nt<T>(&b[m * step]) : b[m * step];
            }
        } else {
#pragma unroll
            for (int m = 0; m < E; ++m) v[m] = mk<T>(0, 0);
        }
        if constexpr (PACK && KIND == K_INVDFT) {
            // Hermitian parts of the two columns through the slot (every
            // lane reaches the barriers; the slot is free before the core)
            const T h = (T)0.5;
            C2 u[E];
            if (packed) {
                const C2* __restrict__ b = in + A.pack_in + (long long)tl * A.in_e;
                const long long step = A.in_e * TPT;
#pragma unroll
                for (int m = 0; m < E; ++m) u[m] = b[m * step];
                Core<T, LOGL, E, INV>::write_owned(v, slot, tl);
            }
            __syncthreads();
            if (packed) {
#pragma unroll
                for (int m = 0; m < E; ++m) {
                    const C2 w = slot[lpad((L - tl - m * TPT) & (L - 1))];
                    v[m] = mk<T>((v[m].x + w.x) * h, (v[m].y - w.y) * h);
                }
            }
            __syncthreads();
            if (packed) Core<T, LOGL, E, INV>::write_owned(u, slot, tl);
            __syncthreads();
            if (packed) {
#pragma unroll
                for (int m = 0; m < E; ++m) {
                    const C2 w = slot[lpad((L - tl - m * TPT) & (L - 1))];
                    const T br = (u[m].x + w.x) * h, bi = (u[m].y - w.y) * h;
                    v[m] = mk<T>(v[m].x - bi, v[m].y + br);
                }
            }
            __syncthreads();
        }
    } else {
        const bool rowm = (A.in_mode == IO_STAGE_ROW);
        const int sc_ = rowm ? c_row : c_cf;
        const int stl = rowm ? tl_row : tl_cf;
        const TransformRef src = (sc_ == c) ? me : locate(A, tile * (unsigned)C + (unsigned)sc_);
        C2* sslot = lds + (long long)sc_ * A.sc;
        if (src.valid) stage_in<T, KIND, L, TPT>(A, src, sslot, stl);
        tile_sync(wv_in);
        const T* rslot = reinterpret_cast<const T*>(slot);
        auto X = [&](int i) -> T { return rslot[ridx(i)]; };
        auto Z = [&](int i) -> C2 { return slot[lpad(i)]; };
        if constexpr (KIND == K_DFT || KIND == K_INVDFT || KIND == K_REALDFT) {
            Core<T, LOGL, E, INV>::read_owned(v, slot, tl);
        } else if constexpr (E >= 2) {
            if constexpr (KIND == K_DCT3 || KIND == K_DST3) t3_to_z<T, KIND, L, E, TPT>(slot, tl, tw, wv);
            pre_owned<T, KIND, L, E, TPT>(v, slot, tl, tw);
        } else {
#pragma unroll
            for (int m = 0; m < E; ++m) {
                const int p = tl + m * TPT;
                v[m] = pre_point<T, KIND>(p, L, tw, X, Z);
            }
        }
        tile_sync(wv);
    }

    // ---------------- FFT core ----------------
    Core<T, LOGL, E, INV>::run(v, slot, tl, tw, wv);

    // two-step twiddle: W_N^{n2*p} on output point p = tl + m*TPT, computed
    // as W_N^{n2*tl} * (W_N^{n2*TPT})^m in FP64 (two table reads instead of E
    // scattered ones; the recurrence adds < 1e-15 relative error)
    if (CPLX && A.fs_logN > 0) {  // (two-step steps are complex passes)
        const unsigned n2 = me.at(A.fs_level);
        const unsigned long long Nfs = 1ull << A.fs_logN;
        const double2* twN = A.tw64 + Nfs;
        double2 w = twN[((unsigned long long)n2 * (unsigned)tl) & (Nfs - 1)];
        const double2 st = twN[((unsigned long long)n2 * (unsigned)TPT) & (Nfs - 1)];
#pragma unroll
        for (int m = 0; m < E; ++m) {
            const C2 wt = mk<T>((T)w.x, (T)w.y);
            v[m] = INV ? cmulc<T>(v[m], wt) : cmul<T>(v[m], wt);
            if (m + 1 < E) w = make_double2(w.x * st.x - w.y * st.y, w.x * st.y + w.y * st.x);
        }
    }

    // ---------------- output side ----------------
    if constexpr (KIND == K_DFT || KIND == K_INVDFT) {
        if (A.out_mode == IO_DIRECT) {
            C2* __restrict__ out = reinterpret_cast<C2*>(A.out) + me.ob;
            if constexpr (PACK && KIND == K_INVDFT) {
                if (packed) {
                    // real parts only: Re IDFT of column 0 and of column H-1
                    const long long es = A.out_e;
                    C2* __restrict__ oa = out + (long long)tl * es;
                    C2* __restrict__ ob = out + A.pack_out + (long long)tl * es;
                    const long long step = es * TPT;
#pragma unroll
                    for (int m = 0; m < E; ++m) {
                        oa[m * step] = mk<T>(v[m].x, (T)0);
                        ob[m * step] = mk<T>(v[m].y, (T)0);
                    }
                    return;
                }
            }
            if constexpr (PACK && KIND == K_DFT) {
                {  // every lane reaches the barrier
                    // Z through the slot (free: the core ends with a barrier
                    // after its last slot read) to pair Z[k] with Z[-k]
                    if (packed) Core<T, LOGL, E, INV>::write_owned(v, slot, tl);
                    __syncthreads();
                    if (packed) {
                        const long long es = A.out_e;
                        C2* __restrict__ oa = out + (long long)tl * es;
                        C2* __restrict__ ob = out + A.pack_out + (long long)tl * es;
                        const long long step = es * TPT;
                        const T h = (T)0.5;
#pragma unroll
                        for (int m = 0; m < E; ++m) {
                            const int p = tl + m * TPT;
                            const C2 z = v[m], w = slot[lpad((L - p) & (L - 1))];
                            oa[m * step] = mk<T>((z.x + w.x) * h, (z.y - w.y) * h);
                            ob[m * step] = mk<T>((z.y + w.y) * h, (w.x - z.x) * h);
                        }
                        return;
                    }
                }
            }
            if (me.valid) {
                if (A.out_e == 1) {
#pragma unroll
                    for (int m = 0; m < E; ++m) {
                        if constexpr (NT) st_nt<T>(&out[tl + m * TPT], v[m]);
                        else out[tl + m * TPT] = v[m];
                    }
                } else {
                    const long long es = A.out_e;
                    C2* __restrict__ b = out + (long long)tl * es;
                    const long long step = es * TPT;
#pragma unroll
                    for (int m = 0; m < E; ++m) {
                        if constexpr (NT) st_nt<T>(&b[m * step], v[m]);
                        else b[m * step] = v[m];
                    }
                }
            }
            return;
        }
    }
    Core<T, LOGL, E, INV>::write_owned(v, slot, tl);
    tile_sync(wv_out);
    {
        const bool rowm = (A.out_mode != IO_STAGE_CF);
        const int sc_ = rowm ? c_row : c_cf;
        const int stl = rowm ? tl_row : tl_cf;
        const TransformRef dst = (sc_ == c) ? me : locate(A, tile * (unsigned)C + (unsigned)sc_);
        const C2* sslot = lds + (long long)sc_ * A.sc;
        if (!dst.valid) return;
        if constexpr (KIND != K_DFT && KIND != K_INVDFT && E >= 2)
            post_real<T, KIND, L, E, TPT>(A, dst, sslot, stl, tw);
        else
            stage_out<T, KIND, L, TPT>(A, dst, sslot, stl, tw);
    }
}

// Completion word of a one-work-group launch, for a host that polls it
// instead of waiting for the stream (the synchronous API's small calls):
// every wavefront makes its stores visible system-wide (L2 written back,
// stores acknowledged), the work-group synchronises, and one lane stores the
// value.  Only launched with a grid of one work-group.
__device__ __forceinline__ void signal_host(const TileArgs& A) {
    __threadfence_system();
    __syncthreads();
    if (threadIdx.x == 0) __hip_atomic_store(A.hflag, A.hval, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

template <typename T, int LOGL, int E, int KIND, bool PACK = false, bool NT = false>
__global__ void __launch_bounds__((tile_threads<T, E>())) tile_kernel(TileArgs A) {
    tile_body<T, LOGL, E, KIND, PACK, NT>(A);
    if (A.hflag) signal_host(A);
}

// ---------------------------------------------------------------------------
// Element-wise kernels for the two-step path of the real kinds and for
// length-1 axes.  One thread per item; same addressing as the tile kernel
// with C = 1 (so b = transform index).
// ---------------------------------------------------------------------------

// pre: kind's input -> complex FFT input (L points per transform) at
// out (contiguous, transform b at b*L)
template <typename T, int KIND>
__global__ void __launch_bounds__(256) pre_kernel(TileArgs A, int L, long long total) {
    using C2 = c2<T>;
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= total) return;
    const unsigned b = (unsigned)(g / L);
    const int p = (int)(g - (long long)b * L);
    const TransformRef r = locate(A, b);
    const C2* __restrict__ tw = reinterpret_cast<const C2*>(A.tw);
    const T* __restrict__ xr = reinterpret_cast<const T*>(A.in);
    const C2* __restrict__ xc = reinterpret_cast<const C2*>(A.in);
    auto X = [&](int i) -> T { return xr[r.ib + (long long)i * A.in_e]; };
    auto Z = [&](int i) -> C2 {
        if constexpr (KIND == K_REALDFT) {
            const long long q = r.ib + 2LL * i * A.in_e;
            return mk<T>(xr[q], xr[q + A.in_e]);
        } else {
            return xc[r.ib + (long long)i * A.in_e];
        }
    };
    C2* __restrict__ out = reinterpret_cast<C2*>(A.out);
    out[(long long)b * L + p] = pre_point<T, KIND>(p, L, tw, X, Z);
}

// post: complex FFT output (contiguous, transform b at b*L) -> kind's output.
// The post-twiddle is computed in FP64 from the FP64 master table whatever
// the data precision: the FFT outputs are widened, combined with exact FP64
// twiddles and rounded once at the store (the FP32 table's rounding would
// otherwise add to the error of these long transforms).
template <typename T, int KIND>
__global__ void __launch_bounds__(256) post_kernel(TileArgs A, int L, long long total) {
    using C2 = c2<T>;
    const int nout = kind_nout(KIND, L);
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= total) return;
    const unsigned b = (unsigned)(g / nout);
    const int i = (int)(g - (long long)b * nout);
    const TransformRef r = locate(A, b);
    const C2* __restrict__ src = reinterpret_cast<const C2*>(A.in) + (long long)b * L;
    auto Tf = [&](int p) -> double2 {
        const C2 v = src[p];
        return make_double2((double)v.x, (double)v.y);
    };
    const double2 yd = post_item<double, KIND>(i, L, A.tw64, Tf);
    const C2 y = mk<T>((T)yd.x, (T)yd.y);
    if constexpr (kind_complex_out(KIND)) {
        reinterpret_cast<C2*>(A.out)[r.ob + (long long)i * A.out_e] = y;
    } else {
        reinterpret_cast<T*>(A.out)[r.ob + (long long)i * A.out_e] = y.x;
    }
}

// length-1 axis: one item per transform
template <typename T, int KIND>
__global__ void __launch_bounds__(256) n1_kernel(TileArgs A) {
    using C2 = c2<T>;
    const unsigned b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= A.ntrans) return;
    const TransformRef r = locate(A, b);
    C2 x;
    if constexpr (kind_complex_in(KIND)) x = reinterpret_cast<const C2*>(A.in)[r.ib];
    else x = mk<T>(reinterpret_cast<const T*>(A.in)[r.ib], (T)0);
    C2 y = n1_item<T, KIND>(x);
    if constexpr (kind_complex_out(KIND)) reinterpret_cast<C2*>(A.out)[r.ob] = y;
    else reinterpret_cast<T*>(A.out)[r.ob] = y.x;
}

}  // namespace mfx
