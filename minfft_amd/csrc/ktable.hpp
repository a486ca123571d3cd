// ktable.hpp -- registry of the compiled kernel instances, filled by the
// per-(precision, kind) instantiation units (inst.hip compiled once per pair).
#pragma once
#include "kinds.hpp"

namespace mfx {

constexpr int kMaxLog = 14;  // largest single-tile FFT length compiled: 2^14

struct KTable {
    const void* tile[K_COUNT][kMaxLog + 1];  // tile_kernel<T, logL, E, kind>
    int tile_E[kMaxLog + 1];                  // points per thread per logL
    // radix-32 variants (32 points per thread, 512-lane work-groups at most):
    // FP32, logL >= 9 only; null elsewhere
    const void* tile32[K_COUNT][kMaxLog + 1];
    // complex tiles with the packed pair (tile_kernel PACK), [inverse][logL]
    const void* tilepk[2][kMaxLog + 1];
    // complex tiles with non-temporal direct I/O (tile_kernel NT), [inverse][logL]
    const void* tilent[2][kMaxLog + 1];
    // small-pass complex tiles, 4 points per lane ([inverse][logL], logL
    // kSmallLogMin..kSmallLogMax; null elsewhere)
    const void* tile4[2][kMaxLog + 1];
    // long real/DCT/DST transforms in two passes (rlong.hpp): step kernels by
    // log2 N1 / log2 N2 (kRlongLogMin..kRlongLogMax; null elsewhere)
    const void* rstep1[K_COUNT][kMaxLog + 1];
    const void* rstep2[K_COUNT][kMaxLog + 1];
    int maxlog;                               // largest logL instantiated
    const void* pre[K_COUNT];
    const void* post[K_COUNT];
    const void* n1[K_COUNT];
};

// lengths of the 4-point-per-lane complex tiles
constexpr int kSmallLogMin = 5;
constexpr int kSmallLogMax = 10;

// maximum single-tile logL compiled per precision
constexpr int kMaxLogF32 = 14;  // 16384 x 8 B = 128 KiB + padding
constexpr int kMaxLogF64 = 13;  //  8192 x 16 B = 128 KiB + padding

void register_f32(KTable& t);
void register_f64(KTable& t);

}  // namespace mfx
