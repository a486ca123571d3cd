// inst.hip -- instantiates the tile, pre, post, two-pass long-transform and
// length-1 kernels of one (precision, kind) pair.  Compiled once per pair with
//   -DINST_F64=0|1 -DINST_KIND=<0..9>
// so the 20 units build in parallel.
#include "tile.hpp"
#include "rlong.hpp"
#include "ktable.hpp"

#ifndef INST_F64
#error "INST_F64 must be defined"
#endif
#ifndef INST_KIND
#error "INST_KIND must be defined"
#endif

namespace mfx {

#if INST_F64
using IT = double;
constexpr int IMAX = kMaxLogF64;
#define REGNAME(k) register_f64_k##k
#else
using IT = float;
constexpr int IMAX = kMaxLogF32;
#define REGNAME(k) register_f32_k##k
#endif

template <int LOGL>
static void reg_tiles(KTable& t) {
    if constexpr (LOGL <= IMAX) {
        constexpr int E = default_E(LOGL);
        t.tile[INST_KIND][LOGL] = reinterpret_cast<const void*>(&tile_kernel<IT, LOGL, E, INST_KIND>);
        t.tile_E[LOGL] = E;
        if constexpr (INST_KIND == K_DFT || INST_KIND == K_INVDFT) {
            t.tilepk[INST_KIND == K_INVDFT][LOGL] =
                reinterpret_cast<const void*>(&tile_kernel<IT, LOGL, E, INST_KIND, true>);
            t.tilent[INST_KIND == K_INVDFT][LOGL] =
                reinterpret_cast<const void*>(&tile_kernel<IT, LOGL, E, INST_KIND, false, true>);
        }
        if constexpr ((INST_KIND == K_DFT || INST_KIND == K_INVDFT) && LOGL >= kSmallLogMin &&
                      LOGL <= kSmallLogMax)
            t.tile4[INST_KIND == K_INVDFT][LOGL] = reinterpret_cast<const void*>(&tile_kernel<IT, LOGL, 4, INST_KIND>);
        reg_tiles<LOGL + 1>(t);
    }
}

#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)
#if INST_F64
#define FNAME CAT(register_f64_k, INST_KIND)
#else
#define FNAME CAT(register_f32_k, INST_KIND)
#endif

// radix-32 tiles (FP32), FFT lengths 2^9 .. 2^IMAX
template <int LOGL>
static void reg_tiles32(KTable& t) {
#if !INST_F64
    if constexpr (LOGL <= IMAX) {
        t.tile32[INST_KIND][LOGL] = reinterpret_cast<const void*>(&tile_kernel<IT, LOGL, 32, INST_KIND>);
        reg_tiles32<LOGL + 1>(t);
    }
#else
    (void)t;
#endif
}

// two-pass step kernels of the long real/DCT/DST transforms
template <int LOG>
static void reg_rlong(KTable& t) {
    if constexpr (INST_KIND != K_DFT && INST_KIND != K_INVDFT && LOG <= kRlongLogMax) {
        t.rstep1[INST_KIND][LOG] = reinterpret_cast<const void*>(&rstep1_kernel<IT, LOG, INST_KIND>);
        t.rstep2[INST_KIND][LOG] = reinterpret_cast<const void*>(&rstep2_kernel<IT, LOG, INST_KIND>);
        reg_rlong<LOG + 1>(t);
    } else {
        (void)t;
    }
}

void FNAME(KTable& t) {
    reg_rlong<kRlongLogMin>(t);
    reg_tiles<0>(t);
    reg_tiles32<9>(t);
    t.maxlog = IMAX;
    t.pre[INST_KIND] = reinterpret_cast<const void*>(&pre_kernel<IT, INST_KIND>);
    t.post[INST_KIND] = reinterpret_cast<const void*>(&post_kernel<IT, INST_KIND>);
    t.n1[INST_KIND] = reinterpret_cast<const void*>(&n1_kernel<IT, INST_KIND>);
}

}  // namespace mfx
