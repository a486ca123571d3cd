// mfma.hpp -- 512-point complex FFT tiles whose radix-16 stage runs on the
// matrix cores as split-FP16 DFT-matrix products (the split-FP16 form of the
// batch extension, MINFFT_SPLIT_FP16).
//
// L = 512 = 16 * 16 * 2, Stockham passes (Ns = 1, 16, 256):
//   pass 1, radix 16 on MFMA: for a group of 16 butterflies (columns j) the
//     16 inputs x_r (rows r) form a 16x16 complex block; X = F16 * x with
//     F16[k][r] = exp(-+2 pi i k r / 16), in real arithmetic
//         Xr = C*xr + S*xi,  Xi = C*xi - S*xr     (forward; S -> -S inverse)
//     C = cos, S = sin of 2 pi k r / 16.  Every operand is split into FP16
//     hi + lo parts (both rounded to nearest) and the product kept to ~22
//     bits as
//         A*x ~ A_hi*x_hi + A_hi*x_lo + A_lo*x_hi
//     with v_mfma_f32_16x16x16f16 accumulating in FP32 (12 MFMAs per
//     group).  Each transform is scaled by a power of two first, so that its
//     largest component lies in [2^13, 2^14): FP16 never overflows and the
//     error stays relative to the transform's own magnitude; the scale is
//     undone exactly on the FP32 result.
//   passes 2 and 3 (radix 16 with twiddles W_256, radix 2 with W_512) run in
//     FP32 on the vector ALUs, as in fft_core.hpp.
// Each wavefront owns two transforms.  Butterflies 2c and 2c+1 (c < 16) form
// the two column groups of a transform, so that lane (c, g) reads the
// adjacent inputs x[2c + 32r], x[2c + 1 + 32r] (r = 4g + q) as one 16-byte
// load straight from memory for contiguous rows; the LDS exchanges after
// pass 1 stay inside the wavefront.
// Operand layouts of v_mfma_f32_16x16x16f16 (lane l, element q = 0..3):
//   A[l%16][4*(l/16)+q], B[4*(l/16)+q][l%16], D[4*(l/16)+q][l%16].
#pragma once
#include "tile.hpp"

namespace mfx {

typedef _Float16 half4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));
typedef _Float16 half2_t __attribute__((ext_vector_type(2)));
typedef float float2_t __attribute__((ext_vector_type(2)));

// 16-byte non-temporal load (two adjacent complex points)
__device__ __forceinline__ float4 ld4_nt(const float2* p) {
    typedef float V4 __attribute__((ext_vector_type(4)));
    const V4 v = __builtin_nontemporal_load(reinterpret_cast<const V4*>(p));
    return make_float4(v.x, v.y, v.z, v.w);
}

__device__ __forceinline__ void split_h(float x, _Float16& hi, _Float16& lo) {
    hi = (_Float16)x;
    lo = (_Float16)(x - (float)hi);
}

// largest value over the 64 lanes of the wavefront (every lane gets it)
__device__ __forceinline__ float wave_max(float m) {
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) m = fmaxf(m, __shfl_xor(m, o));
    return m;
}

// hi + lo FP16 split of four FP32 values (round to nearest for both parts:
// x = hi + lo to ~22 bits), packed as MFMA operands
__device__ __forceinline__ void split4(float a0, float a1, float a2, float a3, half4_t& hi, half4_t& lo) {
    const half2_t h01 = __builtin_convertvector((float2_t){a0, a1}, half2_t);
    const half2_t h23 = __builtin_convertvector((float2_t){a2, a3}, half2_t);
    const half2_t l01 = __builtin_convertvector((float2_t){a0 - (float)h01.x, a1 - (float)h01.y}, half2_t);
    const half2_t l23 = __builtin_convertvector((float2_t){a2 - (float)h23.x, a3 - (float)h23.y}, half2_t);
    hi = __builtin_shufflevector(h01, h23, 0, 1, 2, 3);
    lo = __builtin_shufflevector(l01, l23, 0, 1, 2, 3);
}

// C = A.C transforms per work-group (even, at most 16), 32 lanes each; every
// wavefront owns two whole transforms (2w, 2w+1), so apart from staged
// column I/O the work-group's wavefronts never wait on each other.
// NT: direct I/O with non-temporal accesses (passes larger than the MALL)
template <bool INV, bool NT = false>
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

    // ---- DFT-16 matrix operands A[k][r] = W_16^{kr}, k = cc, r = 4g + q, from
    // the twiddle table (tw[16 + m] = exp(-2 pi i m/16)); read before the
    // inputs so that waiting for them never waits for the inputs ----
    C2 wk[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) wk[q] = twl<float>(tw, 16u + (unsigned)((cc * (4 * g + q)) & 15));

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

    half4_t Ch, Cl, Sh, Sl, NSh, NSl;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        _Float16 h, lo;
        split_h(wk[q].x, h, lo);  // cos
        Ch[q] = h;
        Cl[q] = lo;
        split_h(-wk[q].y, h, lo);  // sin
        Sh[q] = h;
        Sl[q] = lo;
        NSh[q] = -h;
        NSl[q] = -lo;
    }
    // forward: Xr = C xr + S xi, Xi = C xi - S xr; inverse flips S
    const half4_t Pr_h = INV ? NSh : Sh, Pr_l = INV ? NSl : Sl;  // multiplies xi into Xr
    const half4_t Pi_h = INV ? Sh : NSh, Pi_l = INV ? Sl : NSl;  // multiplies xr into Xi

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
            half4_t xrh, xrl, xih, xil;
            split4(xr[0], xr[1], xr[2], xr[3], xrh, xrl);
            split4(xi[0], xi[1], xi[2], xi[3], xih, xil);
            float4_t Xr = {0.f, 0.f, 0.f, 0.f}, Xi = {0.f, 0.f, 0.f, 0.f};
            Xr = __builtin_amdgcn_mfma_f32_16x16x16f16(Cl, xrh, Xr, 0, 0, 0);
            Xr = __builtin_amdgcn_mfma_f32_16x16x16f16(Ch, xrl, Xr, 0, 0, 0);
            Xr = __builtin_amdgcn_mfma_f32_16x16x16f16(Pr_l, xih, Xr, 0, 0, 0);
            Xr = __builtin_amdgcn_mfma_f32_16x16x16f16(Pr_h, xil, Xr, 0, 0, 0);
            Xr = __builtin_amdgcn_mfma_f32_16x16x16f16(Ch, xrh, Xr, 0, 0, 0);
            Xr = __builtin_amdgcn_mfma_f32_16x16x16f16(Pr_h, xih, Xr, 0, 0, 0);
            Xi = __builtin_amdgcn_mfma_f32_16x16x16f16(Cl, xih, Xi, 0, 0, 0);
            Xi = __builtin_amdgcn_mfma_f32_16x16x16f16(Ch, xil, Xi, 0, 0, 0);
            Xi = __builtin_amdgcn_mfma_f32_16x16x16f16(Pi_l, xrh, Xi, 0, 0, 0);
            Xi = __builtin_amdgcn_mfma_f32_16x16x16f16(Pi_h, xrl, Xi, 0, 0, 0);
            Xi = __builtin_amdgcn_mfma_f32_16x16x16f16(Ch, xih, Xi, 0, 0, 0);
            Xi = __builtin_amdgcn_mfma_f32_16x16x16f16(Pi_h, xrh, Xi, 0, 0, 0);
            // position 16(2cc + h) + 4g + q: padded 34cc + 17h + 4g + q
            C2* d = s + (34 * cc + 17 * h + 4 * g);
#pragma unroll
            for (int q = 0; q < 4; ++q) d[q] = make_float2(ldexpf(Xr[q], -sh[tt]), ldexpf(Xi[q], -sh[tt]));
        }
    }
    tile_sync(true);

    // ---- passes 2 and 3 on the vector ALUs, inside the wavefront ----
    C2 v[E];
    Core<float, 9, E, INV>::read_owned(v, slot, tl);
    tile_sync(true);
    stockham_pass<float, L, E, 16, 16, INV, false>(v, slot, tl, tw);
    tile_sync(true);
    Core<float, 9, E, INV>::read_owned(v, slot, tl);
    stockham_pass<float, L, E, 2, 256, INV, true>(v, slot, tl, tw);

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
// butterflies j = w + 8i, i < 4.  The scale is per column (transform c,
// butterfly j): its 16 inputs sit in lanes c + 16g.  Passes 2 and 3 run
// with the c-fast core mapping (lane t: transform t % 16, thread t / 16),
// whose owned positions give coalesced direct stores.
template <bool INV, bool NT = false>
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

    // DFT-16 matrix operands A[k][r] = W_16^{kr}, k = c, r = 4g + q
    C2 wk[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) wk[q] = twl<float>(tw, 16u + (unsigned)((c * (4 * g + q)) & 15));

    // inputs x_c[j + 32(4g + q)] of the wavefront's butterflies j = w + 8i
    C2 x[4][4];
    if (me.valid) {
        // element j + 32(4g + q) = (w + 128g) + 8i + 32q: one base, immediate steps
        const long long es = A.in_e;
        const C2* __restrict__ p0 = reinterpret_cast<const C2*>(A.in) + me.ib + (long long)(w + 128 * g) * es;
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int q = 0; q < 4; ++q)
                x[i][q] = NT ? ld_nt<float>(&p0[(long long)(8 * i + 32 * q) * es]) : p0[(long long)(8 * i + 32 * q) * es];
    } else {
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int q = 0; q < 4; ++q) x[i][q] = make_float2(0.f, 0.f);
    }

    half4_t Ch, Cl, Sh, Sl, NSh, NSl;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        _Float16 h, lo;
        split_h(wk[q].x, h, lo);  // cos
        Ch[q] = h;
        Cl[q] = lo;
        split_h(-wk[q].y, h, lo);  // sin
        Sh[q] = h;
        Sl[q] = lo;
        NSh[q] = -h;
        NSl[q] = -lo;
    }
    const half4_t Pr_h = INV ? NSh : Sh, Pr_l = INV ? NSl : Sl;  // multiplies xi into Xr
    const half4_t Pi_h = INV ? Sh : NSh, Pi_l = INV ? Sl : NSl;  // multiplies xr into Xi

    // ---- pass 1 on the matrix cores: output k = 4g + q of butterfly j of
    // transform c goes to position 16j + k of slot c ----
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int j = w + 8 * i;
        // column scale: largest |component| of the column -> [2^13, 2^14)
        float m = 0.f;
#pragma unroll
        for (int q = 0; q < 4; ++q) m = fmaxf(m, fmaxf(fabsf(x[i][q].x), fabsf(x[i][q].y)));
        m = fmaxf(m, __shfl_xor(m, 16));
        m = fmaxf(m, __shfl_xor(m, 32));
        int ex = 0;
        (void)frexpf(m, &ex);
        const int sh = (m > 0.f) ? 14 - ex : 0;
        half4_t xrh, xrl, xih, xil;
        split4(ldexpf(x[i][0].x, sh), ldexpf(x[i][1].x, sh), ldexpf(x[i][2].x, sh), ldexpf(x[i][3].x, sh), xrh, xrl);
        split4(ldexpf(x[i][0].y, sh), ldexpf(x[i][1].y, sh), ldexpf(x[i][2].y, sh), ldexpf(x[i][3].y, sh), xih, xil);
        float4_t Xr = {0.f, 0.f, 0.f, 0.f}, Xi = {0.f, 0.f, 0.f, 0.f};
        Xr = __builtin_amdgcn_mfma_f32_16x16x16f16(Cl, xrh, Xr, 0, 0, 0);
        Xr = __builtin_amdgcn_mfma_f32_16x16x16f16(Ch, xrl, Xr, 0, 0, 0);
        Xr = __builtin_amdgcn_mfma_f32_16x16x16f16(Pr_l, xih, Xr, 0, 0, 0);
        Xr = __builtin_amdgcn_mfma_f32_16x16x16f16(Pr_h, xil, Xr, 0, 0, 0);
        Xr = __builtin_amdgcn_mfma_f32_16x16x16f16(Ch, xrh, Xr, 0, 0, 0);
        Xr = __builtin_amdgcn_mfma_f32_16x16x16f16(Pr_h, xih, Xr, 0, 0, 0);
        Xi = __builtin_amdgcn_mfma_f32_16x16x16f16(Cl, xih, Xi, 0, 0, 0);
        Xi = __builtin_amdgcn_mfma_f32_16x16x16f16(Ch, xil, Xi, 0, 0, 0);
        Xi = __builtin_amdgcn_mfma_f32_16x16x16f16(Pi_l, xrh, Xi, 0, 0, 0);
        Xi = __builtin_amdgcn_mfma_f32_16x16x16f16(Pi_h, xrl, Xi, 0, 0, 0);
        Xi = __builtin_amdgcn_mfma_f32_16x16x16f16(Ch, xih, Xi, 0, 0, 0);
        Xi = __builtin_amdgcn_mfma_f32_16x16x16f16(Pi_h, xrh, Xi, 0, 0, 0);
        // position 16j + 4g + q: padded 17j + 4g + q
        C2* d = slot + (17 * j + 4 * g);
#pragma unroll
        for (int q = 0; q < 4; ++q) d[q] = make_float2(ldexpf(Xr[q], -sh), ldexpf(Xi[q], -sh));
    }
    __syncthreads();

    // ---- passes 2 and 3 on the vector ALUs, c-fast core ----
    C2 v[E];
    Core<float, 9, E, INV>::read_owned(v, slot, tl);
    __syncthreads();
    stockham_pass<float, L, E, 16, 16, INV, false>(v, slot, tl, tw);
    __syncthreads();
    Core<float, 9, E, INV>::read_owned(v, slot, tl);
    stockham_pass<float, L, E, 2, 256, INV, true>(v, slot, tl, tw);

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

}  // namespace mfx
