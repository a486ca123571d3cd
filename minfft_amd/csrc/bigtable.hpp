// bigtable.hpp -- registry of the kernels outside the per-kind tile table:
// fused two-step, radix-2-split, residue-split, split-FP16 matrix-core and
// the persistent short-length batch kernel.
#pragma once
#include <cstddef>

namespace mfx {

struct BigTable {
    const void* fn[2][17][2];  // [prec 0=f64 1=f32][log2 N][inverse]
    size_t lds[2][17];
    int bt[2][17];
    int wpc[2][17];  // resident work-groups per CU
    // alternative launch shape (smaller work-groups), where instantiated
    const void* alt[2][17][2];
    size_t alt_lds[2][17];
    int alt_bt[2][17];
    int alt_wpc[2][17];
    // 512-point FP32 tiles with the split-FP16 MFMA radix-16 stage (mfma.hpp)
    const void* mfma512[2];
    const void* mfma512cf[2];  // column variant (16 adjacent transforms per work-group)
    const void* mfma512nt[2];  // the two with non-temporal direct I/O
    const void* mfma512cfnt[2];
    // the same with both radix-16 stages on the matrix cores, [column][nt][inverse]
    const void* mfma2[2][2][2];
    // radix-2-split column variant (32 transforms per work-group), [nt][inverse]
    const void* mfma512cs[2][2];
    // single-pass radix-2-split kernels (split2.hpp), by log2 L; variant 1
    // runs the halves with 32 points per lane (radix-32 passes)
    const void* split2[2][17][2][2];  // [prec][log2 L][variant][inverse]
    const void* split2nt[2][17][2][2];  // the same with non-temporal I/O
    int split2_bt[2][17][2];
    size_t split2_lds[2][17];
    // residue-split single-pass kernels (rsplit.hpp), by log2 N, log2 P - 1
    // and residues per work-group - 1
    const void* rsplit[2][17][4][2][2];  // [prec][log2 N][log2 P - 1][B - 1][inverse]
    int rsplit_bt[2][17][4][2];
    size_t rsplit_lds[2][17][4][2];
    // XCD-cooperative two-step kernels (coop.hpp), by log2 N, and the reset
    // kernel of their counters
    const void* coop[2][17][2];  // [prec][log2 N][inverse]
    size_t coop_lds[2][17];
    int coop_bt[2][17];
    int coop_wpc[2][17];
    const void* coop_reset;
    // persistent launch for the short 1D complex groups of a batch (multi.hpp)
    const void* multi[2][2];  // [prec][inverse]
    size_t multi_lds[2];
    const void* multi_reset;
    // radix-2-split column tiles (csplit.hpp), by log2 L
    const void* csplit[2][17][2][2];  // [prec][log2 L][non-temporal][inverse]
    size_t csplit_slot[2][17];        // LDS bytes per transform
    int csplit_tpt[2][17];            // lanes per transform (512 lanes per work-group at most)
};

void register_big(BigTable& t);
void register_rsplit_part0(BigTable& t);  // rsplit_inst.hip, one unit per transform length
void register_rsplit_part1(BigTable& t);
void register_rsplit_part2(BigTable& t);
inline void register_rsplit(BigTable& t) {
    register_rsplit_part0(t);
    register_rsplit_part1(t);
    register_rsplit_part2(t);
}
void register_csplit(BigTable& t);  // csplit_inst.hip

}  // namespace mfx
