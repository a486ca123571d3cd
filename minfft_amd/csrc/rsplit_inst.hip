// rsplit_inst.hip -- instances of the residue-split single-pass kernel
// (rsplit.hpp) and their registry entries in BigTable: lengths past one
// work-group's LDS (and the limit itself), with LDS slots of 2^12..2^14
// points so that one to four work-groups share a CU.
#include "rsplit.hpp"
#include "bigtable.hpp"

namespace mfx {

// <T, log2 M, log2 P, residues per work-group, min waves per SIMD>
template <typename T, int LOGM, int LOGP, int B, int WPE>
static void reg_rsplit(BigTable& t, int prec) {
    constexpr int E = 16;
    const int lg = LOGM + LOGP;
    t.rsplit[prec][lg][LOGP - 1][B - 1][0] = reinterpret_cast<const void*>(&rsplit_kernel<T, LOGM, LOGP, B, E, false, WPE>);
    t.rsplit[prec][lg][LOGP - 1][B - 1][1] = reinterpret_cast<const void*>(&rsplit_kernel<T, LOGM, LOGP, B, E, true, WPE>);
    t.rsplit_bt[prec][lg][LOGP - 1][B - 1] = (1 << LOGM) / E;
    t.rsplit_lds[prec][lg][LOGP - 1][B - 1] = rsplit_lds_bytes<T, LOGM>();
}

// three units (RSPLIT_PART 0, 1, 2: transform lengths 2^14, 2^15, 2^16)
// so that the instances compile in parallel
#ifndef RSPLIT_PART
#error "RSPLIT_PART must be defined"
#endif
#if RSPLIT_PART == 0
void register_rsplit_part0(BigTable& t) {
    // 2^14: 4 x 2^12 (one or two residues per work-group) or 2 x 2^13
    reg_rsplit<float, 12, 2, 1, 4>(t, 1);
    reg_rsplit<float, 12, 2, 2, 4>(t, 1);
    reg_rsplit<float, 13, 1, 1, 4>(t, 1);
    reg_rsplit<double, 12, 2, 1, 2>(t, 0);
    reg_rsplit<double, 12, 2, 2, 2>(t, 0);
}
#elif RSPLIT_PART == 1
void register_rsplit_part1(BigTable& t) {
    // 2^15: 4 x 2^13 or 8 x 2^12
    reg_rsplit<float, 13, 2, 1, 4>(t, 1);
    reg_rsplit<float, 13, 2, 2, 4>(t, 1);
    reg_rsplit<float, 12, 3, 1, 4>(t, 1);
    reg_rsplit<float, 12, 3, 2, 4>(t, 1);
    reg_rsplit<double, 12, 3, 1, 2>(t, 0);
    reg_rsplit<double, 12, 3, 2, 2>(t, 0);
}
#else
void register_rsplit_part2(BigTable& t) {
    // 2^16: 8 x 2^13 (one or two residues per work-group), 4 x 2^14 or
    // 16 x 2^12 (16 blocks with two residues would spill)
    reg_rsplit<float, 13, 3, 1, 4>(t, 1);
    reg_rsplit<float, 13, 3, 2, 4>(t, 1);
    reg_rsplit<float, 14, 2, 1, 4>(t, 1);
    reg_rsplit<float, 14, 2, 2, 4>(t, 1);
    reg_rsplit<float, 12, 4, 1, 4>(t, 1);
    reg_rsplit<double, 12, 4, 1, 2>(t, 0);
}
#endif

}  // namespace mfx
