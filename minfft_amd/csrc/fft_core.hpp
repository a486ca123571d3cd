// fft_core.hpp -- register/LDS Stockham FFT core for gfx950 (CDNA4).
//
// One transform of length L = 2^LOGL is held by TPT = L/E threads, E complex
// points per thread.  Thread tl "owns" positions p = tl + m*TPT (m < E) on
// entry and on exit: inputs are read from, and outputs written to, the same
// owned positions, in natural order.  Between the radix passes the points are
// exchanged through LDS (one slot of the work-group's dynamic LDS per
// transform); the first pass starts from registers and the last pass ends in
// registers, so a plain complex transform touches LDS only L*(passes-1) times.
//
// Pass p (radix R, Ns = product of the radices before it) computes, for
// every butterfly j < L/R with k = j mod Ns:
//     a[r] = x[j + r*L/R] * W_{Ns*R}^{r*k},  a = DFT_R(a),
//     y[(j-k)*R + k + r*Ns] = a[r]
// -- the self-sorting (Stockham) formulation, so no bit reversal is needed.
// Twiddles come from a concatenated table tw[M + m] = exp(-2*pi*i*m/M) for
// every power of two M (built on the device in FP64, see twiddle.hip).
//
// The same arithmetic computes the reference's split-radix DIF result
// (minfft.c:77-272 forward, 293-488 inverse): y_k = sum_n x_n W_N^{+-nk},
// unnormalised.
#pragma once
#include <hip/hip_runtime.h>

namespace mfx {

template <typename T> struct cvec;
template <> struct cvec<float> { using type = float2; };
template <> struct cvec<double> { using type = double2; };
template <typename T> using c2 = typename cvec<T>::type;

template <typename T> __device__ __forceinline__ c2<T> mk(T x, T y) { c2<T> r; r.x = x; r.y = y; return r; }
template <typename T> __device__ __forceinline__ c2<T> add(c2<T> a, c2<T> b) { return mk<T>(a.x + b.x, a.y + b.y); }
template <typename T> __device__ __forceinline__ c2<T> sub(c2<T> a, c2<T> b) { return mk<T>(a.x - b.x, a.y - b.y); }
template <typename T> __device__ __forceinline__ c2<T> neg(c2<T> a) { return mk<T>(-a.x, -a.y); }
template <typename T> __device__ __forceinline__ c2<T> conj(c2<T> a) { return mk<T>(a.x, -a.y); }
// a * w
template <typename T> __device__ __forceinline__ c2<T> cmul(c2<T> a, c2<T> w) {
    return mk<T>(a.x * w.x - a.y * w.y, a.x * w.y + a.y * w.x);
}
// a * conj(w)
template <typename T> __device__ __forceinline__ c2<T> cmulc(c2<T> a, c2<T> w) {
    return mk<T>(a.x * w.x + a.y * w.y, a.y * w.x - a.x * w.y);
}
// a * (-i) and a * (+i)
// Non-temporal (streaming) global load / store of one complex value: for
// data read once and results written once, so that they do not displace
// data that is read again soon (a fused kernel's scratch slots)
template <typename T> struct NtVec;
template <> struct NtVec<float> { typedef float V __attribute__((ext_vector_type(2))); };
template <> struct NtVec<double> { typedef double V __attribute__((ext_vector_type(2))); };
template <typename T>
__device__ __forceinline__ c2<T> ld_nt(const c2<T>* p) {
    typedef typename NtVec<T>::V V;
    const V v = __builtin_nontemporal_load(reinterpret_cast<const V*>(p));
    return mk<T>(v.x, v.y);
}
template <typename T>
__device__ __forceinline__ void st_nt(c2<T>* p, c2<T> a) {
    typedef typename NtVec<T>::V V;
    V v;
    v.x = a.x;
    v.y = a.y;
    __builtin_nontemporal_store(v, reinterpret_cast<V*>(p));
}

template <typename T> __device__ __forceinline__ c2<T> mul_mi(c2<T> a) { return mk<T>(a.y, -a.x); }
template <typename T> __device__ __forceinline__ c2<T> mul_pi(c2<T> a) { return mk<T>(-a.y, a.x); }

constexpr int ilog2(int x) { return x <= 1 ? 0 : 1 + ilog2(x >> 1); }

// Returns x, but hides its value from the optimiser (an empty asm statement
// on a register: no instruction, no memory access).  Used to keep loop-
// invariant twiddle loads inside the loops of the persistent kernels, where
// hoisting them would pin dozens of registers.
__device__ __forceinline__ int opaque(int x) {
    asm volatile("" : "+v"(x));
    return x;
}

// cos(2*pi*k/32), k = 0..8
__device__ constexpr double kCos32[9] = {
    1.0,
    0.9807852804032304491261822,
    0.9238795325112867561281832,
    0.8314696123025452370787884,
    0.7071067811865475244008444,
    0.5555702330196022247428308,
    0.38268343236508977172846,
    0.1950903220161282678482849,
    0.0};

// exp(-2*pi*i*k/32) for the forward direction, its conjugate for the inverse.
// k is a compile-time constant after unrolling; multiples of 45 degrees take
// the exact shortcut.
template <typename T, bool INV>
__device__ __forceinline__ c2<T> rot32(c2<T> a, int k) {
    k &= 31;
    if (INV) k = (32 - k) & 31;  // conjugate rotation
    const T h = (T)0.7071067811865475244008444;
    switch (k) {
        case 0: return a;
        case 4: return mk<T>((a.x + a.y) * h, (a.y - a.x) * h);
        case 8: return mul_mi<T>(a);
        case 12: return mk<T>((a.y - a.x) * h, -(a.x + a.y) * h);
        case 16: return neg<T>(a);
        case 20: return mk<T>(-(a.x + a.y) * h, (a.x - a.y) * h);
        case 24: return mul_pi<T>(a);
        case 28: return mk<T>((a.x - a.y) * h, (a.x + a.y) * h);
        default: break;
    }
    // general angle theta = -2*pi*k/32 = cos - i*sin
    int q = k >> 3, r = k & 7;  // quadrant and offset inside it
    T c = (T)kCos32[r], s = (T)kCos32[8 - r];  // cos/sin of 2*pi*r/32
    // rotate by -(q*90 + r*11.25) degrees: first by the quadrant
    c2<T> b = a;
    if (q == 1) b = mul_mi<T>(a);
    else if (q == 2) b = neg<T>(a);
    else if (q == 3) b = mul_pi<T>(a);
    // then multiply by (c - i s)
    return mk<T>(b.x * c + b.y * s, b.y * c - b.x * s);
}

// In-place radix-R DFT in registers, natural order in and out.
// R in {1,2,4,8,16,32}.  Forward uses exp(-2*pi*i/R); INV the conjugate.
template <int R, bool INV, typename T>
__device__ __forceinline__ void dft(c2<T>* a) {
    if constexpr (R == 1) {
    } else if constexpr (R == 2) {
        c2<T> t = a[0];
        a[0] = add<T>(t, a[1]);
        a[1] = sub<T>(t, a[1]);
    } else if constexpr (R == 4) {
        c2<T> s0 = add<T>(a[0], a[2]), d0 = sub<T>(a[0], a[2]);
        c2<T> s1 = add<T>(a[1], a[3]), d1 = sub<T>(a[1], a[3]);
        c2<T> r1 = INV ? mul_pi<T>(d1) : mul_mi<T>(d1);
        a[0] = add<T>(s0, s1);
        a[2] = sub<T>(s0, s1);
        a[1] = add<T>(d0, r1);
        a[3] = sub<T>(d0, r1);
    } else {
        // R = R1 * R2 Cooley-Tukey: n = R2*n1 + n2, k = k1 + R1*k2
        constexpr int R1 = (R == 8) ? 2 : 4;
        constexpr int R2 = R / R1;
        c2<T> b[R];
#pragma unroll
        for (int n2 = 0; n2 < R2; ++n2) {
            c2<T> t[R1];
#pragma unroll
            for (int n1 = 0; n1 < R1; ++n1) t[n1] = a[R2 * n1 + n2];
            dft<R1, INV, T>(t);
#pragma unroll
            for (int k1 = 0; k1 < R1; ++k1) b[k1 * R2 + n2] = rot32<T, INV>(t[k1], (32 / R) * n2 * k1);
        }
#pragma unroll
        for (int k1 = 0; k1 < R1; ++k1) {
            c2<T> t[R2];
#pragma unroll
            for (int n2 = 0; n2 < R2; ++n2) t[n2] = b[k1 * R2 + n2];
            dft<R2, INV, T>(t);
#pragma unroll
            for (int k2 = 0; k2 < R2; ++k2) a[k1 + R1 * k2] = t[k2];
        }
    }
}

// LDS index with one pad point every 16 to spread the strided pass writes
// over the banks.
__device__ __forceinline__ int lpad(int p) { return p + (p >> 4); }
constexpr int lpad_c(int p) { return p + (p >> 4); }

// Slot length (in complex points) for a transform of length L: room for
// L+1 points (the inverse real transform stages N/2+1 inputs), odd so that
// neighbouring slots start in different banks.
constexpr int slot_len(int L) { return (lpad_c(L) + 1) | 1; }

// Twiddle-table read at an unsigned index: the byte offset is formed in 32
// bits from the (uniform) table base, which lets the compiler use the
// scalar-base + 32-bit vector-offset load form (one address register per
// lane instead of a 64-bit pair).  Tables stay far below 4 GiB.
template <typename T>
__device__ __forceinline__ c2<T> twl(const c2<T>* __restrict__ tw, unsigned i) {
    return *reinterpret_cast<const c2<T>*>(reinterpret_cast<const char*>(tw) + i * (unsigned)sizeof(c2<T>));
}

// Barrier between the LDS write and the LDS read of an exchange.  When every
// transform of the work-group lies inside one wavefront (wave == true), the
// exchange only needs ordering inside the wavefront: a wavefront's LDS
// operations execute in issue order, so a compiler-level fence suffices and
// the wavefronts of the work-group run independently.  Otherwise a
// work-group barrier.  `wave` must be uniform over the work-group.
__device__ __forceinline__ void tile_sync(bool wave) {
    if (wave) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    } else {
        __syncthreads();
    }
}


// Stockham pass twiddles: read every W^{rk} from the table (false) or only
// W^{k}, W^{2k}, W^{4k}, W^{8k} and form the rest as products (true)
#ifndef MINFFT_TW_POW2
#define MINFFT_TW_POW2 1
#endif
constexpr bool kTwiddlePow2Loads = MINFFT_TW_POW2 != 0;

// The R-1 twiddles W^{r k} (r = 1..R-1) of one butterfly from the table
// twr[m] = W^m.  With kTwiddlePow2Loads, only the powers-of-two multiples of
// k are read and the others formed as products of at most three of them: a
// 64-lane read of W^{rk} spans up to r cache lines, so this trades scattered
// reads (bound by the texture path) for a few VALU operations, at a cost of
// a few ulp in the product twiddles.
template <typename T, int R>
__device__ __forceinline__ void butterfly_twiddles(c2<T>* wb, const c2<T>* __restrict__ twr, int k) {
    if constexpr (kTwiddlePow2Loads && R >= 4) {
#pragma unroll
        for (int p2 = 1; p2 < R; p2 *= 2) wb[p2 - 1] = twl<T>(twr, (unsigned)(p2 * k));
#pragma unroll
        for (int r = 3; r < R; ++r) {
            if ((r & (r - 1)) == 0) continue;
            const int hi = 1 << (31 - __builtin_clz(r));  // highest power of two in r
            wb[r - 1] = cmul<T>(wb[hi - 1], wb[r - hi - 1]);
        }
    } else {
#pragma unroll
        for (int r = 1; r < R; ++r) wb[r - 1] = twl<T>(twr, (unsigned)(r * k));
    }
}

template <typename T, int LOGL, int E, bool INV>
struct Core {
    static constexpr int L = 1 << LOGL;
    static_assert(E >= 1 && E <= L, "points per thread");
    static constexpr int TPT = L / E;
    static constexpr int LOGE = ilog2(E);
    static constexpr int NP = (LOGL == 0) ? 0 : (LOGL + LOGE - 1) / LOGE;
    static constexpr int REM = (LOGL == 0) ? 0 : LOGL - (NP - 1) * LOGE;

    // Twiddles of pass P: W_{Ns*R}^{r*k} for r = 1..R-1 of each of the NB
    // butterflies.  Loaded one pass ahead (before the LDS exchange that
    // precedes the pass) so their latency hides behind the exchange.
    template <int P> struct Tw {
        static constexpr int LOGR = (P == 0) ? REM : LOGE;
        static constexpr int LOGNS = (P == 0) ? 0 : REM + (P - 1) * LOGE;
        static constexpr int R = 1 << LOGR;
        static constexpr int Ns = 1 << LOGNS;
        static constexpr int NB = E / R;
        static constexpr int CNT = (Ns > 1) ? NB * (R - 1) : 1;
        c2<T> w[CNT];
        __device__ __forceinline__ void load(int tl, const c2<T>* __restrict__ tw) {
            if constexpr (Ns > 1) {
                const c2<T>* twr = tw + Ns * R;
                const int tlo = opaque(tl);
#pragma unroll
                for (int b = 0; b < NB; ++b) {
                    const int k = (tlo + b * TPT) & (Ns - 1);
                    butterfly_twiddles<T, R>(w + b * (R - 1), twr, k);
                }
            }
        }
    };

    // owned positions tl + m*TPT of the slot <-> registers
    __device__ static __forceinline__ void read_owned(c2<T> (&v)[E], const c2<T>* slot, int tl) {
        if constexpr (TPT % 16 == 0) {
            const c2<T>* sb = slot + lpad(tl);
#pragma unroll
            for (int m = 0; m < E; ++m) v[m] = sb[m * (TPT + TPT / 16)];
        } else {
#pragma unroll
            for (int m = 0; m < E; ++m) v[m] = slot[lpad(tl + m * TPT)];
        }
    }
    __device__ static __forceinline__ void write_owned(const c2<T> (&v)[E], c2<T>* slot, int tl) {
        if constexpr (TPT % 16 == 0) {
            c2<T>* sb = slot + lpad(tl);
#pragma unroll
            for (int m = 0; m < E; ++m) sb[m * (TPT + TPT / 16)] = v[m];
        } else {
#pragma unroll
            for (int m = 0; m < E; ++m) slot[lpad(tl + m * TPT)] = v[m];
        }
    }

    // slot: this transform's LDS slot (complex points, padded by lpad).
    // On entry the slot may be overwritten (caller has synchronised).
    // wave: every transform of the work-group lies in one wavefront (TPT <= 64)
    __device__ static void run(c2<T> (&v)[E], c2<T>* slot, int tl, const c2<T>* __restrict__ tw,
                               bool wave = false) {
        if constexpr (NP > 0) {
            Tw<0> w0;
            pass<0>(v, slot, tl, tw, w0, wave);
        }
    }

    template <int P>
    __device__ static void pass(c2<T> (&v)[E], c2<T>* slot, int tl, const c2<T>* __restrict__ tw,
                                const Tw<P>& twp, bool wave) {
        constexpr int R = Tw<P>::R;
        constexpr int Ns = Tw<P>::Ns;
        constexpr int NB = Tw<P>::NB;
        constexpr bool last = (P == NP - 1);
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            const int j = tl + b * TPT;
            const int k = j & (Ns - 1);
            c2<T> a[R];
#pragma unroll
            for (int r = 0; r < R; ++r) a[r] = v[b + r * NB];
            if constexpr (Ns > 1) {
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    const c2<T> w = twp.w[b * (R - 1) + r - 1];
                    a[r] = INV ? cmulc<T>(a[r], w) : cmul<T>(a[r], w);
                }
            }
            dft<R, INV, T>(a);
            if constexpr (last) {
#pragma unroll
                for (int r = 0; r < R; ++r) v[b + r * NB] = a[r];
            } else {
                const int base = (j - k) * R + k;
                if constexpr (Ns == 1) {
                    // base = R*j: positions base + r, r < R; for R <= 16 they
                    // stay in one 16-point block, for R = 32 the block edge
                    // falls at r = 16 -- both give lpad(base) + r + r/16
                    c2<T>* sb = slot + lpad(base);
#pragma unroll
                    for (int r = 0; r < R; ++r) sb[r + (r >> 4)] = a[r];
                } else if constexpr ((Ns % 16) == 0) {
                    // padded index is linear in r: one base, immediate offsets
                    c2<T>* sb = slot + lpad(base);
                    constexpr int step = Ns + Ns / 16;
#pragma unroll
                    for (int r = 0; r < R; ++r) sb[r * step] = a[r];
                } else {
#pragma unroll
                    for (int r = 0; r < R; ++r) slot[lpad(base + r * Ns)] = a[r];
                }
            }
        }
        if constexpr (!last) {
            Tw<P + 1> nxt;
            nxt.load(tl, tw);
            tile_sync(wave);
            read_owned(v, slot, tl);
            tile_sync(wave);
            pass<P + 1>(v, slot, tl, tw, nxt, wave);
        }
    }
};

// One Stockham pass of radix R (stride Ns) of a length-L transform held as E
// owned points per thread (TPT = L/E threads), outside the Core recursion:
// v holds the owned positions on entry; a non-last pass scatters its results
// into the LDS slot, the last pass leaves them in v at the owned positions.
template <typename T, int L, int E, int R, int Ns, bool INV, bool LAST>
__device__ __forceinline__ void stockham_pass(c2<T> (&v)[E], c2<T>* slot, int tl, const c2<T>* __restrict__ tw) {
    constexpr int TPT = L / E;
    constexpr int NB = E / R;
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        const int j = tl + b * TPT;
        const int k = j & (Ns - 1);
        c2<T> a[R];
#pragma unroll
        for (int r = 0; r < R; ++r) a[r] = v[b + r * NB];
        if constexpr (Ns > 1) {
            c2<T> wv[R - 1];
            butterfly_twiddles<T, R>(wv, tw + Ns * R, k);
#pragma unroll
            for (int r = 1; r < R; ++r) a[r] = INV ? cmulc<T>(a[r], wv[r - 1]) : cmul<T>(a[r], wv[r - 1]);
        }
        dft<R, INV, T>(a);
        if constexpr (LAST) {
#pragma unroll
            for (int r = 0; r < R; ++r) v[b + r * NB] = a[r];
        } else {
            const int base = (j - k) * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) slot[lpad(base + r * Ns)] = a[r];
        }
    }
}

// Points per thread used for a given length: radix-16 passes (E = 16) for
// L >= 16, one pass of radix L below.
constexpr int default_E(int logL) { return logL >= 4 ? 16 : (1 << logL); }

}  // namespace mfx
