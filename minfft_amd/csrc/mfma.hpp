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
//     hi + lo parts and the product kept to ~22 bits as
//         A*x ~ A_hi*x_hi + A_hi*x_lo + A_lo*x_hi
//     with v_mfma_f32_16x16x16f16 accumulating in FP32 (12 MFMAs per
//     group).  Each column is scaled by a power of two first, so that its
//     largest component lies in [2^13, 2^14): the FP16 range is never
//     exceeded and the error stays relative to the column's own magnitude;
//     the scale is undone exactly on the FP32 result.
//   passes 2 and 3 (radix 16 with twiddles W_256, radix 2 with W_512) run in
//     FP32 on the vector ALUs, as in fft_core.hpp.
// Operand layouts of v_mfma_f32_16x16x16f16 (lane l, element q = 0..3):
//   A[l%16][4*(l/16)+q], B[4*(l/16)+q][l%16], D[4*(l/16)+q][l%16].
#pragma once
#include "tile.hpp"

namespace mfx {

typedef _Float16 half4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

// cos / sin of 2*pi*m/16 (m any integer)
__device__ __forceinline__ float cos16(int m) {
    m &= 15;
    // cos(2 pi m/16) = cos(2 pi (2m)/32): fold into [0, 8] of kCos32
    int k = 2 * m;  // in 32nds, 0..30
    if (k > 16) k = 32 - k;                 // cos symmetric
    return k <= 8 ? (float)kCos32[k] : -(float)kCos32[16 - k];
}
__device__ __forceinline__ float sin16(int m) { return cos16(m - 4); }  // sin x = cos(x - pi/2)

__device__ __forceinline__ void split_h(float x, _Float16& hi, _Float16& lo) {
    hi = (_Float16)x;
    lo = (_Float16)(x - (float)hi);
}

template <bool INV>
__global__ void __launch_bounds__(512) mfma512_kernel(TileArgs A) {
    using C2 = float2;
    constexpr int L = 512, E = 16, TPT = L / E;  // 32 threads per transform
    extern __shared__ __align__(16) unsigned char smem_raw[];
    C2* lds = reinterpret_cast<C2*>(smem_raw);
    const C2* __restrict__ tw = reinterpret_cast<const C2*>(A.tw);

    const int t = threadIdx.x;
    const int C = A.C;  // 16 transforms per work-group (512 threads)
    const unsigned tile = tile_id(A);
    const int c_row = t / TPT, tl_row = t - c_row * TPT;
    const int c_cf = t & (C - 1), tl_cf = t >> A.lgC;
    const int c = A.cfast ? c_cf : c_row;
    const int tl = A.cfast ? tl_cf : tl_row;
    const TransformRef me = locate(A, tile * (unsigned)C + (unsigned)c);
    C2* slot = lds + (long long)c * A.sc;

    // ---- input into LDS at natural (padded) positions ----
    if (A.in_mode == IO_DIRECT) {
        const C2* __restrict__ in = reinterpret_cast<const C2*>(A.in) + me.ib;
        C2 v[E];
        if (me.valid) {
            const long long es = A.in_e;
#pragma unroll
            for (int m = 0; m < E; ++m) v[m] = in[(long long)(tl + m * TPT) * es];
        } else {
#pragma unroll
            for (int m = 0; m < E; ++m) v[m] = make_float2(0.f, 0.f);
        }
        Core<float, 9, E, INV>::write_owned(v, slot, tl);
    } else {
        const bool rowm = (A.in_mode == IO_STAGE_ROW);
        const int sc_ = rowm ? c_row : c_cf;
        const int stl = rowm ? tl_row : tl_cf;
        const TransformRef src = (sc_ == c) ? me : locate(A, tile * (unsigned)C + (unsigned)sc_);
        C2* sslot = lds + (long long)sc_ * A.sc;
        if (src.valid) stage_in<float, INV ? K_INVDFT : K_DFT, L, TPT>(A, src, sslot, stl);
    }
    __syncthreads();

    // ---- pass 1: radix 16 on the matrix cores (Ns = 1, no twiddles) ----
    {
        const int lane = t & 63, wave = t >> 6;
        const int col = lane & 15;        // butterfly inside the group
        const int rb = 4 * (lane >> 4);   // first of this lane's 4 rows
        // DFT-matrix operands: A[k][r] with k = lane%16, r = rb + q
        half4_t Ch, Cl, Sh, Sl, NSh, NSl;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int m = col * (rb + q);
            _Float16 h, lo;
            split_h(cos16(m), h, lo);
            Ch[q] = h;
            Cl[q] = lo;
            split_h(sin16(m), h, lo);
            Sh[q] = h;
            Sl[q] = lo;
            NSh[q] = -h;
            NSl[q] = -lo;
        }
        // forward: Xr = C xr + S xi, Xi = C xi - S xr; inverse flips S
        const half4_t Pr_h = INV ? NSh : Sh, Pr_l = INV ? NSl : Sl;  // multiplies xi into Xr
        const half4_t Pi_h = INV ? Sh : NSh, Pi_l = INV ? Sl : NSl;  // multiplies xr into Xi
        // 16 transforms x 2 groups of 16 butterflies = 32 groups over 8 waves
        constexpr int NW = 8, GPW = 4;
        C2 xin[GPW][4];
#pragma unroll
        for (int i = 0; i < GPW; ++i) {
            const int g = wave + i * NW;
            const int tc = g >> 1, j = (g & 1) * 16 + col;
            const C2* s = lds + (long long)tc * A.sc;
#pragma unroll
            for (int q = 0; q < 4; ++q) xin[i][q] = s[lpad(j + (rb + q) * 32)];
        }
        __syncthreads();  // all inputs read before any output is written
#pragma unroll
        for (int i = 0; i < GPW; ++i) {
            const int g = wave + i * NW;
            const int tc = g >> 1, j = (g & 1) * 16 + col;
            // column scale: largest |component| of this column -> [2^13, 2^14)
            float mx = 0.f;
#pragma unroll
            for (int q = 0; q < 4; ++q) mx = fmaxf(mx, fmaxf(fabsf(xin[i][q].x), fabsf(xin[i][q].y)));
            mx = fmaxf(mx, __shfl_xor(mx, 16));
            mx = fmaxf(mx, __shfl_xor(mx, 32));
            int ex = 0;
            (void)frexpf(mx, &ex);
            const int sh = (mx > 0.f) ? 14 - ex : 0;
            half4_t xrh, xrl, xih, xil;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                _Float16 h, lo;
                split_h(ldexpf(xin[i][q].x, sh), h, lo);
                xrh[q] = h;
                xrl[q] = lo;
                split_h(ldexpf(xin[i][q].y, sh), h, lo);
                xih[q] = h;
                xil[q] = lo;
            }
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
            // D[k = rb + q][col]: output k of butterfly j -> position j*16 + k
            C2* s = lds + (long long)tc * A.sc;
#pragma unroll
            for (int q = 0; q < 4; ++q) s[lpad(j * 16 + rb + q)] = make_float2(ldexpf(Xr[q], -sh), ldexpf(Xi[q], -sh));
        }
    }
    __syncthreads();

    // ---- passes 2 and 3 on the vector ALUs ----
    C2 v[E];
    Core<float, 9, E, INV>::read_owned(v, slot, tl);
    __syncthreads();
    stockham_pass<float, L, E, 16, 16, INV, false>(v, slot, tl, tw);
    __syncthreads();
    Core<float, 9, E, INV>::read_owned(v, slot, tl);
    stockham_pass<float, L, E, 2, 256, INV, true>(v, slot, tl, tw);

    // ---- output ----
    if (A.out_mode == IO_DIRECT) {
        if (me.valid) {
            C2* __restrict__ out = reinterpret_cast<C2*>(A.out) + me.ob;
            const long long es = A.out_e;
#pragma unroll
            for (int m = 0; m < E; ++m) out[(long long)(tl + m * TPT) * es] = v[m];
        }
        return;
    }
    __syncthreads();
    Core<float, 9, E, INV>::write_owned(v, slot, tl);
    __syncthreads();
    const bool rowm = (A.out_mode != IO_STAGE_CF);
    const int sc_ = rowm ? c_row : c_cf;
    const int stl = rowm ? tl_row : tl_cf;
    const TransformRef dst = (sc_ == c) ? me : locate(A, tile * (unsigned)C + (unsigned)sc_);
    if (!dst.valid) return;
    const C2* sslot = lds + (long long)sc_ * A.sc;
    stage_out<float, INV ? K_INVDFT : K_DFT, L, TPT>(A, dst, sslot, stl, tw);
}

}  // namespace mfx
