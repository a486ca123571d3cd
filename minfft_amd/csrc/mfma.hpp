// mfma.hpp -- 512-point complex FFT tiles whose radix-16 stage runs on the
// matrix cores as split-FP16 DFT-matrix products (the split-FP16 form of the
// batch extension, MINFFT_SPLIT_FP16).
//
// L = 512 = 16 * 16 * 2, Stockham passes (Ns = 1, 16, 256):
//   pass 1, radix 16 on MFMA (no twiddles): for a group of 16 columns (butterflies, or
//     transforms) the 16 inputs x_r (rows r) form a 16x16 complex block;
//     X = F16 * x with F16[k][r] = exp(-+2 pi i k r / 16), in real arithmetic
//         Xr = C*xr + S*xi,  Xi = C*xi - S*xr     (forward; S -> -S inverse)
//     C = cos, S = sin of 2 pi k r / 16, i.e. Xr = [C | S] [xr ; xi] and
//     Xi = [-S | C] [xr ; xi]: one K = 32 product each.  Every operand is
//     split into FP16 hi + lo parts (both rounded to nearest) and the
//     product kept to ~22 bits as
//         A*x ~ A_lo*x_hi + A_hi*x_lo + A_hi*x_hi
//     with v_mfma_f32_16x16x32_f16 accumulating in FP32 (6 MFMAs per
//     group).  The matrix parts are compile-time constants; the data are
//     scaled by a power of two first, so that the largest component lies in
//     [2^13, 2^14): FP16 never overflows and the error stays relative to the
//     data's own magnitude; the scale is undone exactly on the FP32 result.
//   pass 2, radix 16 with twiddles W_256^{(j % 16) r}: FP32 on the vector
//     ALUs, as in fft_core.hpp; or (STAGES = 2, schedule key mstages=2) the
//     same product on the matrix cores, its inputs twiddled in FP32 first
//     and scaled anew;
//   pass 3, radix 2 with W_512, FP32 on the vector ALUs.
// Operand layouts of v_mfma_f32_16x16x32_f16 (lane l, element j = 0..7,
// q = 0..3): A[l%16][8*(l/16)+j], B[8*(l/16)+j][l%16], D[4*(l/16)+q][l%16].
#pragma once
#include "tile.hpp"

namespace mfx {

typedef _Float16 half4_t __attribute__((ext_vector_type(4)));
typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef float float4_t __attribute__((ext_vector_type(4)));
typedef _Float16 half2_t __attribute__((ext_vector_type(2)));
typedef float float2_t __attribute__((ext_vector_type(2)));

// 16-byte non-temporal load (two adjacent complex points)
__device__ __forceinline__ float4 ld4_nt(const float2* p) {
    typedef float V4 __attribute__((ext_vector_type(4)));
    const V4 v = __builtin_nontemporal_load(reinterpret_cast<const V4*>(p));
    return make_float4(v.x, v.y, v.z, v.w);
}

// DFT-16 matrix operands of every lane of a wavefront for
// v_mfma_f32_16x16x32_f16, split into FP16 hi + lo parts at compile time from
// the exact cosines of multiples of pi/8.  Lane (k, g) = (lane % 16, lane / 16)
// holds A[k][8g + j], j < 8, of the 16x32 blocks Ar = [C | Pr] and
// Ai = [Pi | C] that multiply the 32-row operand [x_r ; x_i]:
//   C  = cos(2 pi k r / 16),
//   Pr = sin (forward) or -sin (inverse) -- multiplies x_i into X_r,
//   Pi = -Pr -- multiplies x_r into X_i.
// Row s = 8g + j of the operand is x_r (g < 2) or x_i (g >= 2) of point
// r = 4h + j + 4 [j >= 4], h = g % 2: the points that lanes (c, h) and
// (c, h + 2) load as r = 4g + q, once their real and imaginary parts are
// swapped between the two halves of the wavefront (mfma32_operand_b).
struct alignas(16) Mfma32Ops {
    _Float16 h[32];  // arh[8] arl[8] aih[8] ail[8]
};
struct Mfma32Table {
    Mfma32Ops e[2][64];  // [inverse][lane]
};
constexpr double kCosPi8[16] = {1.0, 0.92387953251128675613, 0.70710678118654752440, 0.38268343236508977173,
                                0.0, -0.38268343236508977173, -0.70710678118654752440, -0.92387953251128675613,
                                -1.0, -0.92387953251128675613, -0.70710678118654752440, -0.38268343236508977173,
                                0.0, 0.38268343236508977173, 0.70710678118654752440, 0.92387953251128675613};
constexpr Mfma32Table make_mfma32_table() {
    Mfma32Table t{};
    for (int inv = 0; inv < 2; ++inv)
        for (int lane = 0; lane < 64; ++lane) {
            const int k = lane & 15, g = lane >> 4;
            for (int j = 0; j < 8; ++j) {
                const int r = 4 * (g & 1) + j + (j >= 4 ? 4 : 0);
                const int m = (k * r) & 15;
                const double c = kCosPi8[m], sn = kCosPi8[(m + 12) & 15];  // cos, sin of pi m / 8
                const double pr = inv ? -sn : sn;
                const double ar = g < 2 ? c : pr, ai = g < 2 ? -pr : c;
                const _Float16 arh = (_Float16)ar, aih = (_Float16)ai;
                _Float16* h = t.e[inv][lane].h;
                h[j] = arh;
                h[8 + j] = (_Float16)(ar - (double)arh);
                h[16 + j] = aih;
                h[24 + j] = (_Float16)(ai - (double)aih);
            }
        }
    return t;
}
__device__ const Mfma32Table kMfma32 = make_mfma32_table();

struct Mfma32Operands {
    half8_t arh, arl, aih, ail;
};
// this lane's operands (four 16-byte loads of an 8 KiB table)
template <bool INV>
__device__ __forceinline__ Mfma32Operands mfma32_operands(int lane) {
    const half8_t* p = reinterpret_cast<const half8_t*>(kMfma32.e[INV ? 1 : 0][lane].h);
    return Mfma32Operands{p[0], p[1], p[2], p[3]};
}

// largest value over the 64 lanes of the wavefront (every lane gets it)
__device__ __forceinline__ float wave_max(float m) {
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) m = fmaxf(m, __shfl_xor(m, o));
    return m;
}

// -1.0f in a register the compiler does not see as a constant, so that
// fma(hi, m1, x) stays an fma with an FP16 operand (one v_fma_mix_f32)
// instead of being folded into a conversion and a subtraction
__device__ __forceinline__ float opaque_m1() {
    float m1;
    asm volatile("s_mov_b32 %0, -1.0" : "=s"(m1));
    return m1;
}

// hi + lo FP16 split of four FP32 values (round to nearest for both parts:
// x = hi + lo to ~22 bits), packed as MFMA operands; m1 = opaque_m1()
__device__ __forceinline__ void split4(float a0, float a1, float a2, float a3, float m1, half4_t& hi, half4_t& lo) {
    const half2_t h01 = __builtin_convertvector((float2_t){a0, a1}, half2_t);
    const half2_t h23 = __builtin_convertvector((float2_t){a2, a3}, half2_t);
    // x - hi is exact in FP32; one rounding to FP16
    const half2_t l01 = __builtin_convertvector(
        (float2_t){__builtin_fmaf((float)h01.x, m1, a0), __builtin_fmaf((float)h01.y, m1, a1)}, half2_t);
    const half2_t l23 = __builtin_convertvector(
        (float2_t){__builtin_fmaf((float)h23.x, m1, a2), __builtin_fmaf((float)h23.y, m1, a3)}, half2_t);
    hi = __builtin_shufflevector(h01, h23, 0, 1, 2, 3);
    lo = __builtin_shufflevector(l01, l23, 0, 1, 2, 3);
}

// The 32-row operand [x_r ; x_i] of a column (hi and lo parts) from lane
// (c, g)'s points r = 4g + q: lanes of the lower half of the wavefront keep
// their real parts and take those of lane + 32, lanes of the upper half keep
// their imaginary parts and take those of lane - 32 (v_permlane32_swap)
__device__ __forceinline__ void mfma32_operand_b(const float (&xr)[4], const float (&xi)[4], float m1, half8_t& bh,
                                                 half8_t& bl) {
    float u[4], v[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const auto sw = __builtin_amdgcn_permlane32_swap(__float_as_uint(xr[q]), __float_as_uint(xi[q]), false, false);
        u[q] = __uint_as_float(sw[0]);
        v[q] = __uint_as_float(sw[1]);
    }
    half4_t uh, ul, vh, vl;
    split4(u[0], u[1], u[2], u[3], m1, uh, ul);
    split4(v[0], v[1], v[2], v[3], m1, vh, vl);
    bh = __builtin_shufflevector(uh, vh, 0, 1, 2, 3, 4, 5, 6, 7);
    bl = __builtin_shufflevector(ul, vl, 0, 1, 2, 3, 4, 5, 6, 7);
}

// A*b to ~22 bits from hi/lo parts (lo*lo dropped), smallest terms first
__device__ __forceinline__ float4_t mfma32_product(half8_t ah, half8_t al, half8_t bh, half8_t bl) {
    float4_t d = {0.f, 0.f, 0.f, 0.f};
    d = __builtin_amdgcn_mfma_f32_16x16x32_f16(al, bh, d, 0, 0, 0);
    d = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bl, d, 0, 0, 0);
    d = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bh, d, 0, 0, 0);
    return d;
}

// Power-of-two exponent that brings m (a largest |component|) into
// [2^13, 2^14): FP16 never overflows and the split error stays relative to
// the data's own magnitude
__device__ __forceinline__ int scale_exp(float m) {
    int ex = 0;
    (void)frexpf(m, &ex);
    return (m > 0.f) ? 14 - ex : 0;
}

// One 16-column group of a radix-16 stage: lane (c, g) holds inputs
// r = 4g + q (q < 4) of column c in x; scaled by 2^sh, DFT-16 on the matrix
// cores, unscaled; output k = 4g + q of column c returned in y[q]
__device__ __forceinline__ void mfma16_columns(const float2 (&x)[4], int sh, const Mfma32Operands& op, float m1,
                                               float2 (&y)[4]) {
    float xr[4], xi[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        xr[q] = ldexpf(x[q].x, sh);
        xi[q] = ldexpf(x[q].y, sh);
    }
    half8_t bh, bl;
    mfma32_operand_b(xr, xi, m1, bh, bl);
    const float4_t Xr = mfma32_product(op.arh, op.arl, bh, bl);
    const float4_t Xi = mfma32_product(op.aih, op.ail, bh, bl);
#pragma unroll
    for (int q = 0; q < 4; ++q) y[q] = make_float2(ldexpf(Xr[q], -sh), ldexpf(Xi[q], -sh));
}

// C = A.C transforms per work-group (even, at most 16), 32 lanes each; every
// wavefront owns two whole transforms (2w, 2w+1), so apart from staged
// column I/O the work-group's wavefronts never wait on each other.
// Butterflies 2c and 2c+1 (c < 16) form the two column groups of a
// transform, so that lane (c, g) reads the adjacent inputs x[2c + 32r],
// x[2c + 1 + 32r] (r = 4g + q) as one 16-byte load straight from memory for
// contiguous rows; the LDS exchanges after pass 1 stay inside the wavefront.
// NT: direct I/O with non-temporal accesses (passes larger than the MALL)
template <bool INV, bool NT = false, int STAGES = 1>
__global__ void __launch_bounds__(512) mfma512_kernel(TileArgs A) {
    using C2 = float2;
    constexpr int L = 512, E = 16, TPT = L / E;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    C2* lds = reinterpret_cast<C2*>(smem_raw);
    const C2* __restrict__ tw = reinterpret_cast<const C2*>(A.tw);

    const int t = threadIdx.x;
    const int C = A.C;
    const unsigned tile = tile_id(A);
    const int c = t >> 5, tl = t & (TPT - 1);  // row-mapped core
    const int lane = t & 63, w = t >> 6;
    const int cc = lane & 15, g = lane >> 4;    // MFMA operand coordinates
    const TransformRef me = locate(A, tile * (unsigned)C + (unsigned)c);
    C2* slot = lds + (long long)c * A.sc;
    C2* wslot = lds + (long long)(2 * w) * A.sc;  // the wavefront's first slot

    // ---- DFT-16 matrix operands from the compile-time table, read before
    // the inputs so that waiting for them never waits for the inputs ----
    const Mfma32Operands op = mfma32_operands<INV>(lane);
    const float m1 = opaque_m1();

    // ---- pass-1 inputs in MFMA operand layout: lane (cc, g) holds, for each
    // of the wavefront's transforms tt and q < 4, the pair
    // (x[2cc + 32r], x[2cc + 1 + 32r]), r = 4g + q: input r of butterflies
    // 2cc (h = 0: .x .y) and 2cc + 1 (h = 1: .z .w)
    float4 P[2][4];
    if (A.in_mode == IO_DIRECT) {
        // contiguous rows (item stride 1), straight from memory
#pragma unroll
        for (int tt = 0; tt < 2; ++tt) {
            const unsigned lo = __builtin_amdgcn_readlane((unsigned)(unsigned long long)me.ib, 32 * tt);
            const unsigned hi = __builtin_amdgcn_readlane((unsigned)((unsigned long long)me.ib >> 32), 32 * tt);
            const long long ib = (long long)(((unsigned long long)hi << 32) | lo);
            const bool ok = __builtin_amdgcn_readlane((int)me.valid, 32 * tt) != 0;
            const C2* __restrict__ row = reinterpret_cast<const C2*>(A.in) + ib;
            const C2* __restrict__ src = row + 2 * cc + 128 * g;
            if (!ok) {
#pragma unroll
                for (int q = 0; q < 4; ++q) P[tt][q] = make_float4(0.f, 0.f, 0.f, 0.f);
            } else if ((reinterpret_cast<unsigned long long>(row) & 15) == 0) {
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    P[tt][q] = NT ? ld4_nt(src + 32 * q)
                                  : *static_cast<const float4*>(__builtin_assume_aligned(src + 32 * q, 16));
            } else {
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const C2 a = NT ? ld_nt<float>(src + 32 * q) : src[32 * q];
                    const C2 b = NT ? ld_nt<float>(src + 32 * q + 1) : src[32 * q + 1];
                    P[tt][q] = make_float4(a.x, a.y, b.x, b.y);
                }
            }
        }
    } else {
        const bool rowm = (A.in_mode == IO_STAGE_ROW);
        const int c_cf = t & (C - 1), tl_cf = t >> A.lgC;
        const int sc_ = rowm ? c : c_cf;
        const int stl = rowm ? tl : tl_cf;
        const TransformRef src = (sc_ == c) ? me : locate(A, tile * (unsigned)C + (unsigned)sc_);
        C2* sslot = lds + (long long)sc_ * A.sc;
        if (src.valid) stage_in<float, INV ? K_INVDFT : K_DFT, L, TPT>(A, src, sslot, stl);
        tile_sync(rowm);  // row staging: a wavefront stages its own transforms
        // (slots of transforms past the batch end hold stale data: their
        // results are never stored)
#pragma unroll
        for (int tt = 0; tt < 2; ++tt) {
            const C2* s = wslot + (long long)tt * A.sc;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                // lpad(2cc + 32r) and lpad(2cc + 1 + 32r) are adjacent
                const C2* pp = s + lpad(2 * cc + 32 * (4 * g + q));
                P[tt][q] = make_float4(pp[0].x, pp[0].y, pp[1].x, pp[1].y);
            }
        }
    }

    // ---- per-transform power-of-two scale: largest |component| -> [2^13, 2^14),
    // so FP16 never overflows and the split error stays relative to the
    // transform's own magnitude
    int sh[2];
#pragma unroll
    for (int tt = 0; tt < 2; ++tt) {
        float m = 0.f;
#pragma unroll
        for (int q = 0; q < 4; ++q)
            m = fmaxf(fmaxf(m, fmaxf(fabsf(P[tt][q].x), fabsf(P[tt][q].y))),
                      fmaxf(fabsf(P[tt][q].z), fabsf(P[tt][q].w)));
        m = wave_max(m);
        int ex = 0;
        (void)frexpf(m, &ex);
        sh[tt] = __builtin_amdgcn_readfirstlane((m > 0.f) ? 14 - ex : 0);
    }

    // forward: Xr = C xr + S xi, Xi = C xi - S xr; inverse flips S
    // ---- pass 1: radix 16 on the matrix cores (Ns = 1, no twiddles); output
    // k = 4g + q of butterfly j = 2cc + h goes to position 16j + k ----
#pragma unroll
    for (int tt = 0; tt < 2; ++tt) {
        C2* s = wslot + (long long)tt * A.sc;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            float xr[4], xi[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                xr[q] = ldexpf(h ? P[tt][q].z : P[tt][q].x, sh[tt]);
                xi[q] = ldexpf(h ? P[tt][q].w : P[tt][q].y, sh[tt]);
            }
            half8_t bh, bl;
            mfma32_operand_b(xr, xi, m1, bh, bl);
            const float4_t Xr = mfma32_product(op.arh, op.arl, bh, bl);
            const float4_t Xi = mfma32_product(op.aih, op.ail, bh, bl);
            // position 16(2cc + h) + 4g + q: padded 34cc + 17h + 4g + q
            C2* d = s + (34 * cc + 17 * h + 4 * g);
#pragma unroll
            for (int q = 0; q < 4; ++q) d[q] = make_float2(ldexpf(Xr[q], -sh[tt]), ldexpf(Xi[q], -sh[tt]));
        }
    }
    tile_sync(true);

    C2 v[E];
    if constexpr (STAGES == 2) {
        // ---- pass 2 (radix 16, twiddles W_256^{(j % 16) r}) on the matrix
        // cores: inputs x[j + 32r], r = 4g + q, of butterflies j = 2cc + h from
        // the wavefront's slots, twiddled in FP32, one scale per transform;
        // output k goes to 256 (j / 16) + j % 16 + 16k ----
        {
            float2 x2[2][2][4];
#pragma unroll
            for (int tt = 0; tt < 2; ++tt) {
                const C2* s = wslot + (long long)tt * A.sc;
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const int j = 2 * cc + h;
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const int r = 4 * g + q;
                        const C2 a = s[j + (j >> 4) + 34 * r];  // lpad(j + 32r)
                        const C2 wt = tw[256 + (j & 15) * r];
                        x2[tt][h][q] = INV ? cmulc<float>(a, wt) : cmul<float>(a, wt);
                    }
                }
            }
            tile_sync(true);  // every input read before the slots are overwritten
#pragma unroll
            for (int tt = 0; tt < 2; ++tt) {
                float m = 0.f;
#pragma unroll
                for (int h = 0; h < 2; ++h)
#pragma unroll
                    for (int q = 0; q < 4; ++q) m = fmaxf(m, fmaxf(fabsf(x2[tt][h][q].x), fabsf(x2[tt][h][q].y)));
                const int sh2 = __builtin_amdgcn_readfirstlane(scale_exp(wave_max(m)));
                C2* s = wslot + (long long)tt * A.sc;
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const int j = 2 * cc + h;
                    float2 y[4];
                    mfma16_columns(x2[tt][h], sh2, op, m1, y);
                    // lpad(256 (j >> 4) + (j & 15) + 16k) = 272 (j >> 4) + (j & 15) + 17k
                    C2* d = s + (272 * (j >> 4) + (j & 15));
#pragma unroll
                    for (int q = 0; q < 4; ++q) d[17 * (4 * g + q)] = y[q];
                }
            }
        }
        tile_sync(true);

        // ---- pass 3 (radix 2, W_512) on the vector ALUs, inside the wavefront ----
        Core<float, 9, E, INV>::read_owned(v, slot, tl);
        stockham_pass<float, L, E, 2, 256, INV, true>(v, slot, tl, tw);
    } else {
        // ---- passes 2 and 3 on the vector ALUs, inside the wavefront ----
        Core<float, 9, E, INV>::read_owned(v, slot, tl);
        tile_sync(true);
        stockham_pass<float, L, E, 16, 16, INV, false>(v, slot, tl, tw);
        tile_sync(true);
        Core<float, 9, E, INV>::read_owned(v, slot, tl);
        stockham_pass<float, L, E, 2, 256, INV, true>(v, slot, tl, tw);
    }

    // ---- output ----
    if (A.out_mode == IO_DIRECT) {
        if (me.valid) {
            const Walk<C2> wo(reinterpret_cast<C2*>(A.out) + me.ob, 0, 1, tl, TPT, A.out_e);
#pragma unroll
            for (int m = 0; m < E; ++m) {
                if constexpr (NT) st_nt<float>(&wo[m], v[m]);
                else wo[m] = v[m];
            }
        }
        return;
    }
    tile_sync(true);
    Core<float, 9, E, INV>::write_owned(v, slot, tl);
    const bool rowm = (A.out_mode != IO_STAGE_CF);
    tile_sync(rowm);
    const int c_cf = t & (C - 1), tl_cf = t >> A.lgC;
    const int sc_ = rowm ? c : c_cf;
    const int stl = rowm ? tl : tl_cf;
    const TransformRef dst = (sc_ == c) ? me : locate(A, tile * (unsigned)C + (unsigned)sc_);
    if (!dst.valid) return;
    const C2* sslot = lds + (long long)sc_ * A.sc;
    stage_out<float, INV ? K_INVDFT : K_DFT, L, TPT>(A, dst, sslot, stl, tw);
}

// Column variant for strided transforms whose neighbours are adjacent in
// memory (item stride e, batch stride 1 on both sides): 16 transforms per
// work-group (512 lanes), and the MFMA columns are the 16 transforms rather
// than 16 butterflies.  Group (butterfly j, transforms c < 16): lane (c, g)
// reads x_c[j + 32r], r = 4g + q -- for fixed (g, q) the 16 lanes read 16
// adjacent points, 128 contiguous bytes -- and wavefront w takes the
// butterflies j = 4w + i, i < 4.  One scale serves the wavefront's four
// columns of transform c (their 64 inputs sit in lanes c + 16g).  Passes 2
// and 3 run
// with the c-fast core mapping (lane t: transform t % 16, thread t / 16),
// whose owned positions give coalesced direct stores.
template <bool INV, bool NT = false, int STAGES = 1>
__global__ void __launch_bounds__(512) mfma512_cf_kernel(TileArgs A) {
    using C2 = float2;
    constexpr int L = 512, E = 16, TPT = L / E, CT = 16;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    C2* lds = reinterpret_cast<C2*>(smem_raw);
    const C2* __restrict__ tw = reinterpret_cast<const C2*>(A.tw);

    const int t = threadIdx.x;
    const unsigned tile = tile_id(A);
    const int c = t & (CT - 1), tl = t >> 4;  // c-fast core: transform c, thread tl
    const int g = (t >> 4) & 3, w = t >> 6;     // MFMA operand row block, wavefront
    const TransformRef me = locate(A, tile * (unsigned)CT + (unsigned)c);
    C2* slot = lds + (long long)c * A.sc;

    // DFT-16 matrix operands from the compile-time table
    const Mfma32Operands op = mfma32_operands<INV>(t & 63);
    const float m1 = opaque_m1();

    // inputs x_c[j + 32(4g + q)] of the wavefront's butterflies j = 4w + i
    C2 x[4][4];
    if (me.valid) {
        // element j + 32(4g + q) = (4w + 128g) + i + 32q: one base, immediate steps
        const long long es = A.in_e;
        const C2* __restrict__ p0 = reinterpret_cast<const C2*>(A.in) + me.ib + (long long)(4 * w + 128 * g) * es;
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int q = 0; q < 4; ++q)
                x[i][q] = NT ? ld_nt<float>(&p0[(long long)(i + 32 * q) * es]) : p0[(long long)(i + 32 * q) * es];
    } else {
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int q = 0; q < 4; ++q) x[i][q] = make_float2(0.f, 0.f);
    }

    // ---- pass 1 on the matrix cores: output k = 4g + q of butterfly j of
    // transform c goes to position 16j + k of slot c.  One scale for the
    // wavefront's four columns of transform c (lanes c + 16g): their largest
    // |component| ----
    auto group_scale = [&]() {
        float m = 0.f;
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int q = 0; q < 4; ++q) m = fmaxf(m, fmaxf(fabsf(x[i][q].x), fabsf(x[i][q].y)));
        m = fmaxf(m, __shfl_xor(m, 16));
        m = fmaxf(m, __shfl_xor(m, 32));
        return scale_exp(m);
    };
    {
        const int sh = group_scale();
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int j = 4 * w + i;
            float2 y[4];
            mfma16_columns(x[i], sh, op, m1, y);
            // position 16j + 4g + q: padded 17j + 4g + q
            C2* d = slot + (17 * j + 4 * g);
#pragma unroll
            for (int q = 0; q < 4; ++q) d[q] = y[q];
        }
    }
    __syncthreads();

    C2 v[E];
    if constexpr (STAGES == 2) {
        // ---- pass 2 (radix 16, twiddles W_256^{(j % 16) r}) on the matrix
        // cores: inputs x_c[j + 32r], r = 4g + q, of butterflies j = 4w + i from
        // the slot, twiddled in FP32; output k goes to 256 (j / 16) + j % 16 + 16k ----
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int j = 4 * w + i;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int r = 4 * g + q;
                const C2 a = slot[j + (j >> 4) + 34 * r];  // lpad(j + 32r)
                const C2 wt = tw[256 + (j & 15) * r];
                x[i][q] = INV ? cmulc<float>(a, wt) : cmul<float>(a, wt);
            }
        }
        __syncthreads();  // every input read before the slot is overwritten
        {
            const int sh = group_scale();
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int j = 4 * w + i;
                float2 y[4];
                mfma16_columns(x[i], sh, op, m1, y);
                // lpad(256 (j >> 4) + (j & 15) + 16k) = 272 (j >> 4) + (j & 15) + 17k
                C2* d = slot + (272 * (j >> 4) + (j & 15));
#pragma unroll
                for (int q = 0; q < 4; ++q) d[17 * (4 * g + q)] = y[q];
            }
        }
        __syncthreads();

        // ---- pass 3 (radix 2, W_512) on the vector ALUs, c-fast core ----
        Core<float, 9, E, INV>::read_owned(v, slot, tl);
        stockham_pass<float, L, E, 2, 256, INV, true>(v, slot, tl, tw);
    } else {
        // ---- passes 2 and 3 on the vector ALUs, c-fast core ----
        Core<float, 9, E, INV>::read_owned(v, slot, tl);
        __syncthreads();
        stockham_pass<float, L, E, 16, 16, INV, false>(v, slot, tl, tw);
        __syncthreads();
        Core<float, 9, E, INV>::read_owned(v, slot, tl);
        stockham_pass<float, L, E, 2, 256, INV, true>(v, slot, tl, tw);
    }

    // ---- output: neighbouring lanes store neighbouring transforms ----
    if (me.valid) {
        const Walk<C2> wo(reinterpret_cast<C2*>(A.out) + me.ob, 0, 1, tl, TPT, A.out_e);
#pragma unroll
        for (int mm = 0; mm < E; ++mm) {
            if constexpr (NT) st_nt<float>(&wo[mm], v[mm]);
            else wo[mm] = v[mm];
        }
    }
}

// Radix-2-split column variant (the split column tile of csplit.hpp with the
// radix-16 stage of each half on the matrix cores): 32 adjacent transforms
// per work-group (256-byte row segments of FP32 columns), two work-groups
// per CU.  a[n] = x[n] + x[n+256], b[n] = (x[n] - x[n+256]) W_512^n in
// registers; each 256-point half then runs its first radix-16 pass
// (butterflies j < 16 over rows j + 16r, output 16j + k) as split-FP16
// products and its second (radix 16, W_256) on the vector ALUs; X[2k] and
// X[2k+1] go to neighbouring rows.  Wavefront w takes transforms
// 16 (w & 1) + c and butterflies 4 (w >> 1) + i for the matrix-core pass, so
// the two wavefronts of a pair read the two 128-byte halves of the same
// rows; the vector pass and the stores use the c-fast mapping over all 32.
template <bool INV, bool NT = false>
__global__ void __launch_bounds__(512, 4) mfma512_cs_kernel(TileArgs A) {
    using C2 = float2;
    constexpr int H = 256, E = 16, TPT = H / E, CT = 32;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    C2* lds = reinterpret_cast<C2*>(smem_raw);
    const C2* __restrict__ tw = reinterpret_cast<const C2*>(A.tw);

    const int t = threadIdx.x;
    const unsigned tile = tile_id(A);
    const int w = t >> 6, g = (t >> 4) & 3;
    const int cm = 16 * (w & 1) + (t & 15);  // matrix-core pass: transform
    const int jb = 4 * (w >> 1);             // and its first butterfly
    const int cc = t & (CT - 1), tl = t >> 5;  // vector pass: c-fast core
    const TransformRef mm = locate(A, tile * (unsigned)CT + (unsigned)cm);
    const TransformRef mc = locate(A, tile * (unsigned)CT + (unsigned)cc);
    C2* slot_m = lds + (long long)cm * A.sc;
    C2* slot_c = lds + (long long)cc * A.sc;

    // inputs n = j + 16r (j = jb + i, r = 4g + q) and n + 256
    C2 a[4][4], b[4][4];
    if (mm.valid) {
        const long long es = A.in_e;
        const C2* __restrict__ p0 = reinterpret_cast<const C2*>(A.in) + mm.ib + (long long)(jb + 64 * g) * es;
        const long long half = 256 * es;
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const long long o = (long long)(i + 16 * q) * es;
                const C2 x0 = NT ? ld_nt<float>(&p0[o]) : p0[o];
                const C2 x1 = NT ? ld_nt<float>(&p0[o + half]) : p0[o + half];
                const C2 e = twl<float>(tw, (unsigned)(512 + jb + 64 * g + i + 16 * q));  // W_512^n
                const C2 d = sub<float>(x0, x1);
                a[i][q] = add<float>(x0, x1);
                b[i][q] = INV ? cmulc<float>(d, e) : cmul<float>(d, e);
            }
    } else {
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int q = 0; q < 4; ++q) a[i][q] = b[i][q] = make_float2(0.f, 0.f);
    }
    C2* __restrict__ out = reinterpret_cast<C2*>(A.out) + mc.ob;
    const long long os = A.out_e;
#pragma unroll
    for (int hf = 0; hf < 2; ++hf) {
        C2 (&x)[4][4] = hf ? b : a;
        // ---- pass 1 on the matrix cores; one scale per transform over the
        // wavefront's four butterflies of it (lanes c + 16g).  The operands
        // are loaded per half (an opaque lane index keeps the loads here):
        // held across the vector pass they would push the second half's
        // inputs out of the 128-register budget ----
        {
            const Mfma32Operands op = mfma32_operands<INV>(opaque(t & 63));
            const float m1 = opaque_m1();
            float m = 0.f;
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int q = 0; q < 4; ++q) m = fmaxf(m, fmaxf(fabsf(x[i][q].x), fabsf(x[i][q].y)));
            m = fmaxf(m, __shfl_xor(m, 16));
            m = fmaxf(m, __shfl_xor(m, 32));
            const int sh = scale_exp(m);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int j = jb + i;
                float2 y[4];
                mfma16_columns(x[i], sh, op, m1, y);
                C2* d = slot_m + (17 * j + 4 * g);  // lpad(16j + 4g + q)
#pragma unroll
                for (int q = 0; q < 4; ++q) d[q] = y[q];
            }
        }
        __syncthreads();
        // ---- pass 2 (radix 16, W_256) on the vector ALUs ----
        C2 v[E];
        Core<float, 8, E, INV>::read_owned(v, slot_c, tl);
        stockham_pass<float, H, E, 16, 16, INV, true>(v, slot_c, tl, tw);  // v: DFT at k = tl + m*TPT
        if (mc.valid) {
#pragma unroll
            for (int m = 0; m < E; ++m) {
                const long long row = 2LL * (tl + m * TPT) + hf;
                if constexpr (NT) st_nt<float>(&out[row * os], v[m]);
                else out[row * os] = v[m];
            }
        }
        __syncthreads();  // the slot is read before the next half's pass 1 writes it
    }
}

}  // namespace mfx
