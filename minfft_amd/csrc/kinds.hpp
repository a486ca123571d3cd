// kinds.hpp -- per-point formulas that turn the real, DCT and DST transforms
// into a half-length complex FFT (L = N/2), shared by the fused tile kernel
// (readers over LDS) and the stand-alone pre/post kernels used for lengths
// that need the two-step FFT (readers over global memory).
//
// Each formula follows the reference's reduction step for step:
//   R2C  post-twiddle          minfft.c:547-578
//   C2R  pre-twiddle           minfft.c:628-656
//   DCT-2/DST-2 reorder, post  minfft.c:707-725, 756-774
//   DCT-3/DST-3 pre, reorder   minfft.c:807-824, 857-874
//   DCT-4/DST-4 pre, post      minfft.c:906-921, 953-968
// Twiddles: tw[M + m] = exp(-2*pi*i*m/M); with N = 2L the reference's
// tables are  R2C/C2R e[n] = W_N^n -> tw[2L+n];  T2/T3 e[n] = W_{4N}^n ->
// tw[8L+n];  T4 e1[n] = W_{2N}^n -> tw[4L+n],  e2[m] = W_{8N}^{2m+1} ->
// tw[16L+2m+1].
#pragma once
#include "fft_core.hpp"

namespace mfx {

enum Kind : int {
    K_DFT = 0, K_INVDFT = 1, K_REALDFT = 2, K_INVREALDFT = 3,
    K_DCT2 = 4, K_DST2 = 5, K_DCT3 = 6, K_DST3 = 7, K_DCT4 = 8, K_DST4 = 9,
    K_COUNT = 10
};

__host__ __device__ constexpr bool kind_inverse_fft(int k) {
    return k == K_INVDFT || k == K_INVREALDFT || k == K_DCT3 || k == K_DST3;
}
__host__ __device__ constexpr bool kind_complex_in(int k) {
    return k == K_DFT || k == K_INVDFT || k == K_INVREALDFT;
}
__host__ __device__ constexpr bool kind_complex_out(int k) {
    return k == K_DFT || k == K_INVDFT || k == K_REALDFT;
}
// number of input / output items per transform whose FFT length is L
// (complex items for complex sides, real items for real sides)
__host__ __device__ constexpr int kind_nin(int k, int L) {
    return (k == K_DFT || k == K_INVDFT || k == K_REALDFT) ? L : (k == K_INVREALDFT ? L + 1 : 2 * L);
}
__host__ __device__ constexpr int kind_nout(int k, int L) {
    return (k == K_DFT || k == K_INVDFT) ? L : (k == K_REALDFT ? L + 1 : 2 * L);
}

template <typename T> __device__ __forceinline__ T sqrt2_v() { return (T)1.414213562373095048801688724209698079; }

// ---------------------------------------------------------------------------
// pre-processing: FFT input point p (0 <= p < L) from the transform's input.
// RI(i) returns real input item i; CI(i) complex input item i.
// ---------------------------------------------------------------------------

// inverse real DFT pre-twiddle, reading z[0..L] through Z(i).  The twiddle
// source tw is indexed as the table (tw[M + m] = W_M^m): a pointer to it or
// an accessor with the same operator[].
template <typename T, class TWT, class ZF>
__device__ __forceinline__ c2<T> c2r_pre(int p, int L, TWT tw, ZF Z) {
    if (p == 0) {
        // imaginary parts of z[0] and z[L] are ignored (minfft.c:629-630)
        T a = Z(0).x, b = Z(L).x;
        return mk<T>(a + b, a - b);
    }
    const int h = L >> 1;
    if (p == h) {
        c2<T> z = Z(h);
        return mk<T>(2 * z.x, -2 * z.y);
    }
    const int m = (p < h) ? p : L - p;
    c2<T> a = Z(m), b = Z(L - m);
    c2<T> e = tw[2 * L + m];
    T ur = a.x + b.x, ui = a.y - b.y;
    T ttr = a.x - b.x, tti = a.y + b.y;
    // te = i*conj(e)
    T ter = e.y, tei = e.x;
    T vr = ttr * ter - tti * tei;
    T vi = ttr * tei + tti * ter;
    if (p < h) return mk<T>(ur + vr, ui + vi);
    return mk<T>(ur - vr, -ui + vi);
}

// real DFT post-twiddle: output item k (0 <= k <= L) from FFT output T(i)
template <typename T, class TWT, class TF>
__device__ __forceinline__ c2<T> r2c_post(int k, int L, TWT tw, TF Tf) {
    if (k == 0) {
        c2<T> u = Tf(0);
        return mk<T>(u.x + u.y, (T)0);
    }
    if (k == L) {
        c2<T> u = Tf(0);
        return mk<T>(u.x - u.y, (T)0);
    }
    const int h = L >> 1;
    if (k == h) return conj<T>(Tf(h));
    const int m = (k < h) ? k : L - k;
    c2<T> a = Tf(m), b = Tf(L - m);
    c2<T> e = tw[2 * L + m];
    T ur = (a.x + b.x) / 2, ui = (a.y - b.y) / 2;
    T ttr = a.x - b.x, tti = a.y + b.y;
    // te = e / i
    T ter = e.y, tei = -e.x;
    T vr = (ttr * ter - tti * tei) / 2;
    T vi = (ttr * tei + tti * ter) / 2;
    if (k < h) return mk<T>(ur + vr, ui + vi);
    return mk<T>(ur - vr, -ui + vi);
}

// DCT-2/DST-2 input reorder: t[n] = x[2n] (n < L), t[L+n] = s*x[N-1-2n]
template <typename T, bool DST, class RF>
__device__ __forceinline__ T t2_reorder(int i, int L, RF X) {
    if (i < L) return X(2 * i);
    T v = X(4 * L - 1 - 2 * i);  // N-1-2(i-L) with N = 2L
    return DST ? -v : v;
}

// DCT-3/DST-3 intermediate z[k] (0 <= k <= L) from the real input
template <typename T, bool DST, class TWT, class RF>
__device__ __forceinline__ c2<T> t3_z(int k, int L, TWT tw, RF X) {
    const int N = 2 * L;
    if (k == 0) return mk<T>(DST ? X(N - 1) : X(0), (T)0);
    if (k == L) return mk<T>(sqrt2_v<T>() * (DST ? X(L - 1) : X(L)), (T)0);
    c2<T> e = tw[4 * N + k];
    T xa = DST ? X(N - k - 1) : X(k);
    T xb = DST ? X(k - 1) : X(N - k);
    return mk<T>(xa * e.x - xb * e.y, -xa * e.y - xb * e.x);
}

// FFT input point p for every kind
template <typename T, int KIND, class TWT, class RF, class CF>
__device__ __forceinline__ c2<T> pre_point(int p, int L, TWT tw, RF X, CF Z) {
    if constexpr (KIND == K_DFT || KIND == K_INVDFT || KIND == K_REALDFT) {
        return Z(p);
    } else if constexpr (KIND == K_INVREALDFT) {
        return c2r_pre<T>(p, L, tw, Z);
    } else if constexpr (KIND == K_DCT2 || KIND == K_DST2) {
        constexpr bool D = (KIND == K_DST2);
        return mk<T>(t2_reorder<T, D>(2 * p, L, X), t2_reorder<T, D>(2 * p + 1, L, X));
    } else if constexpr (KIND == K_DCT3 || KIND == K_DST3) {
        constexpr bool D = (KIND == K_DST3);
        auto zf = [&](int k) { return t3_z<T, D>(k, L, tw, X); };
        return c2r_pre<T>(p, L, tw, zf);
    } else {  // DCT-4 / DST-4
        const int N = 2 * L;
        c2<T> e = tw[2 * N + p];
        T xa = X(2 * p), xb = X(N - 1 - 2 * p);
        if constexpr (KIND == K_DCT4)
            return mk<T>(e.x * xa - e.y * xb, e.x * xb + e.y * xa);
        else
            return mk<T>(-e.x * xa - e.y * xb, e.x * xb - e.y * xa);
    }
}

// DCT-2/DST-2: both output items anchored at FFT position n (0 <= n < L)
// from one real-DFT value Zf(n) and one twiddle -- the formulas of
// post_item's DCT-2/DST-2 branch for items n and N - n (DCT-2) or n - 1 and
// N - 1 - n (DST-2); for n = 0 the items 0 and L (DCT-2) or N - 1 and L - 1
// (DST-2), which use Zf(0) and Zf(L).
template <typename T, bool DST, class TWT, class TF>
__device__ __forceinline__ void t2_post_pair(int n, int L, TWT tw, TF Tf, T& a, T& b) {
    const int N = 2 * L;
    if (n == 0) {
        a = 2 * r2c_post<T>(0, L, tw, Tf).x;
        b = sqrt2_v<T>() * r2c_post<T>(L, L, tw, Tf).x;
        return;
    }
    const c2<T> z = r2c_post<T>(n, L, tw, Tf);
    const c2<T> e = tw[4 * N + n];
    const T re = 2 * (z.x * e.x - z.y * e.y), im = -2 * (z.x * e.y + z.y * e.x);
    a = DST ? im : re;  // DCT-2 item n (re) / DST-2 item n - 1 (re = false)
    b = DST ? re : im;  // DCT-2 item N - n / DST-2 item N - 1 - n
}

// ---------------------------------------------------------------------------
// post-processing: output item i from the FFT output T(p), p < L.
// Complex kinds return the complex item; real kinds return .x only.
// ---------------------------------------------------------------------------
template <typename T, int KIND, class TWT, class TF>
__device__ __forceinline__ c2<T> post_item(int i, int L, TWT tw, TF Tf) {
    if constexpr (KIND == K_DFT || KIND == K_INVDFT) {
        return Tf(i);
    } else if constexpr (KIND == K_REALDFT) {
        return r2c_post<T>(i, L, tw, Tf);
    } else if constexpr (KIND == K_INVREALDFT) {
        c2<T> w = Tf(i >> 1);
        return mk<T>((i & 1) ? w.y : w.x, (T)0);
    } else if constexpr (KIND == K_DCT2 || KIND == K_DST2) {
        const int N = 2 * L;
        auto Zf = [&](int k) { return r2c_post<T>(k, L, tw, Tf); };
        int n;
        bool re;
        T v;
        if constexpr (KIND == K_DCT2) {
            if (i == 0) return mk<T>(2 * Zf(0).x, (T)0);
            if (i == L) return mk<T>(sqrt2_v<T>() * Zf(L).x, (T)0);
            if (i < L) { n = i; re = true; } else { n = N - i; re = false; }
        } else {
            if (i == N - 1) return mk<T>(2 * Zf(0).x, (T)0);
            if (i == L - 1) return mk<T>(sqrt2_v<T>() * Zf(L).x, (T)0);
            if (i < L - 1) { n = i + 1; re = false; } else { n = N - 1 - i; re = true; }
        }
        c2<T> z = Zf(n);
        c2<T> e = tw[4 * N + n];
        if (re) v = 2 * (z.x * e.x - z.y * e.y);
        else v = -2 * (z.x * e.y + z.y * e.x);
        return mk<T>(v, (T)0);
    } else if constexpr (KIND == K_DCT3 || KIND == K_DST3) {
        const int N = 2 * L;
        int src;
        bool negate = false;
        if ((i & 1) == 0) src = i >> 1;
        else { src = L + ((N - 1 - i) >> 1); negate = (KIND == K_DST3); }
        c2<T> w = Tf(src >> 1);
        T v = (src & 1) ? w.y : w.x;
        return mk<T>(negate ? -v : v, (T)0);
    } else {  // DCT-4 / DST-4
        const int N = 2 * L;
        const int n = i >> 1;
        T v;
        if ((i & 1) == 0) {
            c2<T> e = tw[8 * N + 4 * n + 1];
            c2<T> t = Tf(n);
            if constexpr (KIND == K_DCT4) v = 2 * (e.x * t.x - e.y * t.y);
            else v = 2 * (e.x * t.y + e.y * t.x);
        } else {
            c2<T> e = tw[8 * N + 4 * n + 3];
            c2<T> t = Tf(L - 1 - n);
            if constexpr (KIND == K_DCT4) v = 2 * (e.x * t.x + e.y * t.y);
            else v = 2 * (-e.x * t.y + e.y * t.x);
        }
        return mk<T>(v, (T)0);
    }
}

// Length-1 transforms along an axis (minfft.c:519-523, 607-610, 700-704,
// 749-753, 798-802, 848-852, 897-901, 944-948).
template <typename T, int KIND>
__device__ __forceinline__ c2<T> n1_item(c2<T> x) {
    if constexpr (KIND == K_DFT || KIND == K_INVDFT) return x;
    else if constexpr (KIND == K_REALDFT) return mk<T>(x.x, (T)0);
    else if constexpr (KIND == K_INVREALDFT) return mk<T>(x.x, (T)0);
    else if constexpr (KIND == K_DCT2 || KIND == K_DST2) return mk<T>(2 * x.x, (T)0);
    else if constexpr (KIND == K_DCT3 || KIND == K_DST3) return mk<T>(x.x, (T)0);
    else return mk<T>(sqrt2_v<T>() * x.x, (T)0);
}

}  // namespace mfx
