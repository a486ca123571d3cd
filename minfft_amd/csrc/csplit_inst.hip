// csplit_inst.hip -- instances of the radix-2-split column tile
// (csplit.hpp) and their registry entries in BigTable: FP32 columns of 512
// and 1024 points (32 and 16 columns of 256 and 128 points per work-group,
// two work-groups per CU), FP64 columns of 512 points (16 columns).
#include "csplit.hpp"
#include "bigtable.hpp"

namespace mfx {

// <T, log2 H, points per lane, min waves per SIMD>
template <typename T, int LOGH, int E, int WPE>
static void reg_csplit(BigTable& t, int prec) {
    const int lg = LOGH + 1;
    t.csplit[prec][lg][0][0] = reinterpret_cast<const void*>(&csplit_kernel<T, LOGH, E, false, WPE, false>);
    t.csplit[prec][lg][0][1] = reinterpret_cast<const void*>(&csplit_kernel<T, LOGH, E, true, WPE, false>);
    t.csplit[prec][lg][1][0] = reinterpret_cast<const void*>(&csplit_kernel<T, LOGH, E, false, WPE, true>);
    t.csplit[prec][lg][1][1] = reinterpret_cast<const void*>(&csplit_kernel<T, LOGH, E, true, WPE, true>);
    t.csplit_slot[prec][lg] = csplit_slot_bytes<T, LOGH>();
    t.csplit_tpt[prec][lg] = (1 << LOGH) / E;
}

void register_csplit(BigTable& t) {
    reg_csplit<float, 8, 16, 4>(t, 1);
    reg_csplit<float, 9, 16, 4>(t, 1);
    reg_csplit<double, 8, 16, 2>(t, 0);
}

}  // namespace mfx
