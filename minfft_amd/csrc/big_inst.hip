// big_inst.hip -- instances of the kernels outside the per-kind tile table,
// and their registry (bigtable.hpp): the fused two-step kernel (big.hpp; FP32
// with 512-thread work-groups, FP64 with 256: two work-groups of ~70 KiB LDS
// per CU where the registers allow), the radix-2-split single-pass kernel
// (split2.hpp), the split-FP16 matrix-core 512-point tile (mfma.hpp) and the
// persistent short-length batch kernel (multi.hpp).
#include "big.hpp"
#include "coop.hpp"
#include "mfma.hpp"
#include "split2.hpp"
#include "multi.hpp"
#include "bigtable.hpp"

namespace mfx {

template <typename T, int L1, int L2, int BT, int WPE>
static void reg_one(BigTable& t, int prec, int variant = 0) {
    const int lg = L1 + L2;
    if (variant) {
        t.alt[prec][lg][0] = reinterpret_cast<const void*>(&big_kernel<T, L1, L2, false, BT, WPE>);
        t.alt[prec][lg][1] = reinterpret_cast<const void*>(&big_kernel<T, L1, L2, true, BT, WPE>);
        t.alt_lds[prec][lg] = big_lds_bytes<T, L1, L2, BT>();
        t.alt_bt[prec][lg] = BT;
        t.alt_wpc[prec][lg] = (4 * WPE * 64) / BT;
        return;
    }
    t.fn[prec][lg][0] = reinterpret_cast<const void*>(&big_kernel<T, L1, L2, false, BT, WPE>);
    t.fn[prec][lg][1] = reinterpret_cast<const void*>(&big_kernel<T, L1, L2, true, BT, WPE>);
    t.lds[prec][lg] = big_lds_bytes<T, L1, L2, BT>();
    t.bt[prec][lg] = BT;
    t.wpc[prec][lg] = (4 * WPE * 64) / BT;  // work-groups per CU the registers allow
}

__global__ void coop_reset_kernel(unsigned* ctr, unsigned n) {
    const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) ctr[i] = 0u;
}

template <typename T, int L1, int L2, int BT, int WPE>
static void reg_coop(BigTable& t, int prec) {
    const int lg = L1 + L2;
    t.coop[prec][lg][0] = reinterpret_cast<const void*>(&coop_kernel<T, L1, L2, false, BT, WPE>);
    t.coop[prec][lg][1] = reinterpret_cast<const void*>(&coop_kernel<T, L1, L2, true, BT, WPE>);
    t.coop_lds[prec][lg] = coop_lds_bytes<T, L1, L2, BT>();
    t.coop_bt[prec][lg] = BT;
    t.coop_wpc[prec][lg] = (4 * WPE * 64) / BT;
}

// variant 0: 16 points per lane; variant 1 (E = 32): 32 points per lane
template <typename T, int LOGH, int WPE, int E = 16>
static void reg_split2(BigTable& t, int prec) {
    constexpr int v = E == 32 ? 1 : 0;
    t.split2[prec][LOGH + 1][v][0] = reinterpret_cast<const void*>(&split2_kernel<T, LOGH, E, false, WPE, false>);
    t.split2[prec][LOGH + 1][v][1] = reinterpret_cast<const void*>(&split2_kernel<T, LOGH, E, true, WPE, false>);
    t.split2nt[prec][LOGH + 1][v][0] = reinterpret_cast<const void*>(&split2_kernel<T, LOGH, E, false, WPE, true>);
    t.split2nt[prec][LOGH + 1][v][1] = reinterpret_cast<const void*>(&split2_kernel<T, LOGH, E, true, WPE, true>);
    t.split2_bt[prec][LOGH + 1][v] = (1 << LOGH) / E;
    t.split2_lds[prec][LOGH + 1] = split2_lds_bytes<T, LOGH>();
}

void register_big(BigTable& t) {
    t.multi[0][0] = reinterpret_cast<const void*>(&multi_kernel<double, false>);
    t.multi[0][1] = reinterpret_cast<const void*>(&multi_kernel<double, true>);
    t.multi[1][0] = reinterpret_cast<const void*>(&multi_kernel<float, false>);
    t.multi[1][1] = reinterpret_cast<const void*>(&multi_kernel<float, true>);
    t.multi_lds[0] = multi_lds_bytes<double>();
    t.multi_lds[1] = multi_lds_bytes<float>();
    t.multi_reset = reinterpret_cast<const void*>(&multi_reset_kernel<unsigned>);
    t.mfma512[0] = reinterpret_cast<const void*>(&mfma512_kernel<false>);
    t.mfma512[1] = reinterpret_cast<const void*>(&mfma512_kernel<true>);
    t.mfma512cf[0] = reinterpret_cast<const void*>(&mfma512_cf_kernel<false>);
    t.mfma512cf[1] = reinterpret_cast<const void*>(&mfma512_cf_kernel<true>);
    t.mfma512nt[0] = reinterpret_cast<const void*>(&mfma512_kernel<false, true>);
    t.mfma512nt[1] = reinterpret_cast<const void*>(&mfma512_kernel<true, true>);
    t.mfma512cfnt[0] = reinterpret_cast<const void*>(&mfma512_cf_kernel<false, true>);
    t.mfma512cfnt[1] = reinterpret_cast<const void*>(&mfma512_cf_kernel<true, true>);
    t.mfma512cs[0][0] = reinterpret_cast<const void*>(&mfma512_cs_kernel<false, false>);
    t.mfma512cs[0][1] = reinterpret_cast<const void*>(&mfma512_cs_kernel<true, false>);
    t.mfma512cs[1][0] = reinterpret_cast<const void*>(&mfma512_cs_kernel<false, true>);
    t.mfma512cs[1][1] = reinterpret_cast<const void*>(&mfma512_cs_kernel<true, true>);
    t.mfma2[0][0][0] = reinterpret_cast<const void*>(&mfma512_kernel<false, false, 2>);
    t.mfma2[0][0][1] = reinterpret_cast<const void*>(&mfma512_kernel<true, false, 2>);
    t.mfma2[0][1][0] = reinterpret_cast<const void*>(&mfma512_kernel<false, true, 2>);
    t.mfma2[0][1][1] = reinterpret_cast<const void*>(&mfma512_kernel<true, true, 2>);
    t.mfma2[1][0][0] = reinterpret_cast<const void*>(&mfma512_cf_kernel<false, false, 2>);
    t.mfma2[1][0][1] = reinterpret_cast<const void*>(&mfma512_cf_kernel<true, false, 2>);
    t.mfma2[1][1][0] = reinterpret_cast<const void*>(&mfma512_cf_kernel<false, true, 2>);
    t.mfma2[1][1][1] = reinterpret_cast<const void*>(&mfma512_cf_kernel<true, true, 2>);
    // <T, log2 H, min waves per SIMD>: the first length past one
    // work-group's LDS, and the two lengths below it (two work-groups per CU)
    reg_split2<float, 14, 4>(t, 1);
    reg_split2<float, 13, 4>(t, 1);
    reg_split2<float, 12, 3>(t, 1);
    reg_split2<double, 13, 2>(t, 0);
    reg_split2<double, 12, 2>(t, 0);
    reg_split2<double, 11, 2>(t, 0);
    reg_split2<float, 14, 2, 32>(t, 1);
    reg_split2<float, 13, 2, 32>(t, 1);
    // <T, log2 N1, log2 N2, threads, min waves per SIMD>
    reg_one<float, 7, 7, 512, 2>(t, 1);
    // FP32 2^15 as 2^7 x 2^8: the 2^8 x 2^7 split spills at this register
    // budget (352 bytes of scratch per lane), this one does not (123 VGPRs)
    reg_one<float, 7, 8, 512, 4>(t, 1);
    reg_one<float, 8, 8, 512, 4>(t, 1);
    reg_one<double, 7, 7, 256, 2>(t, 0);
    reg_one<double, 8, 7, 256, 2>(t, 0);
    reg_one<double, 8, 8, 256, 2>(t, 0);
    // XCD-cooperative forms of the same factorisations
    t.coop_reset = reinterpret_cast<const void*>(&coop_reset_kernel);
    reg_coop<float, 7, 8, 512, 4>(t, 1);  // the fused kernel's 2^15 split (bit-identical results)
    reg_coop<float, 8, 8, 512, 4>(t, 1);
    reg_coop<double, 7, 7, 256, 2>(t, 0);
    reg_coop<double, 8, 7, 256, 2>(t, 0);
    reg_coop<double, 8, 8, 256, 2>(t, 0);
    // smaller work-groups (4 per CU) for launch-shape comparison
    reg_one<float, 8, 7, 256, 4>(t, 1, 1);
    reg_one<float, 8, 8, 256, 4>(t, 1, 1);
}

}  // namespace mfx
