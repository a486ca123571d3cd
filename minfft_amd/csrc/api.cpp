// api.cpp -- the 27 minfft.h entry points over the engine.  Compiled twice:
// as is for libminfft.so (double) and with -DMINFFT_SINGLE=1 for
// libminfft_single.so (float), mirroring the reference's MINFFT_SINGLE
// switch (minfft.h:16-29).
#include "minfft.h"

#include "engine.hpp"

#if MINFFT_SINGLE
static constexpr mfx::Prec kPrec = mfx::PREC_F32;
#else
static constexpr mfx::Prec kPrec = mfx::PREC_F64;
#endif

static mfx::Plan* plan_of(const minfft_aux* a, int kind) {
    mfx::Plan* p = reinterpret_cast<mfx::Plan*>(const_cast<minfft_aux*>(a));
    if (!p || p->magic != 0x6d667870u) {
        mfx::set_error("invalid aux data");
        return nullptr;
    }
    if (p->prec != kPrec) {
        mfx::set_error("aux data was made by the other precision's library");
        return nullptr;
    }
    if (mfx::family_of(kind) != p->fam) {
        mfx::set_error("aux data does not match the transform kind");
        return nullptr;
    }
    return p;
}

static void run(int kind, const void* x, void* y, const minfft_aux* a) {
    mfx::Plan* p = plan_of(a, kind);
    if (p) (void)mfx::execute(*p, kind, x, y);
}

static minfft_aux* mk(mfx::Family f, int d, const int* Ns) {
    return reinterpret_cast<minfft_aux*>(mfx::make_plan(f, kPrec, d, Ns));
}

extern "C" {

void minfft_dft(minfft_cmpl* x, minfft_cmpl* y, const minfft_aux* a) { run(mfx::K_DFT, x, y, a); }
void minfft_invdft(minfft_cmpl* x, minfft_cmpl* y, const minfft_aux* a) { run(mfx::K_INVDFT, x, y, a); }
void minfft_realdft(minfft_real* x, minfft_cmpl* z, const minfft_aux* a) { run(mfx::K_REALDFT, x, z, a); }
void minfft_invrealdft(minfft_cmpl* z, minfft_real* y, const minfft_aux* a) { run(mfx::K_INVREALDFT, z, y, a); }
void minfft_dct2(minfft_real* x, minfft_real* y, const minfft_aux* a) { run(mfx::K_DCT2, x, y, a); }
void minfft_dst2(minfft_real* x, minfft_real* y, const minfft_aux* a) { run(mfx::K_DST2, x, y, a); }
void minfft_dct3(minfft_real* x, minfft_real* y, const minfft_aux* a) { run(mfx::K_DCT3, x, y, a); }
void minfft_dst3(minfft_real* x, minfft_real* y, const minfft_aux* a) { run(mfx::K_DST3, x, y, a); }
void minfft_dct4(minfft_real* x, minfft_real* y, const minfft_aux* a) { run(mfx::K_DCT4, x, y, a); }
void minfft_dst4(minfft_real* x, minfft_real* y, const minfft_aux* a) { run(mfx::K_DST4, x, y, a); }

minfft_aux* minfft_mkaux_dft(int d, int* Ns) { return mk(mfx::FAM_DFT, d, Ns); }
minfft_aux* minfft_mkaux_dft_1d(int N) { int n[1] = {N}; return mk(mfx::FAM_DFT, 1, n); }
minfft_aux* minfft_mkaux_dft_2d(int N1, int N2) { int n[2] = {N1, N2}; return mk(mfx::FAM_DFT, 2, n); }
minfft_aux* minfft_mkaux_dft_3d(int N1, int N2, int N3) { int n[3] = {N1, N2, N3}; return mk(mfx::FAM_DFT, 3, n); }
minfft_aux* minfft_mkaux_realdft(int d, int* Ns) { return mk(mfx::FAM_REALDFT, d, Ns); }
minfft_aux* minfft_mkaux_realdft_1d(int N) { int n[1] = {N}; return mk(mfx::FAM_REALDFT, 1, n); }
minfft_aux* minfft_mkaux_realdft_2d(int N1, int N2) { int n[2] = {N1, N2}; return mk(mfx::FAM_REALDFT, 2, n); }
minfft_aux* minfft_mkaux_realdft_3d(int N1, int N2, int N3) { int n[3] = {N1, N2, N3}; return mk(mfx::FAM_REALDFT, 3, n); }
minfft_aux* minfft_mkaux_t2t3(int d, int* Ns) { return mk(mfx::FAM_T2T3, d, Ns); }
minfft_aux* minfft_mkaux_t2t3_1d(int N) { int n[1] = {N}; return mk(mfx::FAM_T2T3, 1, n); }
minfft_aux* minfft_mkaux_t2t3_2d(int N1, int N2) { int n[2] = {N1, N2}; return mk(mfx::FAM_T2T3, 2, n); }
minfft_aux* minfft_mkaux_t2t3_3d(int N1, int N2, int N3) { int n[3] = {N1, N2, N3}; return mk(mfx::FAM_T2T3, 3, n); }
minfft_aux* minfft_mkaux_t4(int d, int* Ns) { return mk(mfx::FAM_T4, d, Ns); }
minfft_aux* minfft_mkaux_t4_1d(int N) { int n[1] = {N}; return mk(mfx::FAM_T4, 1, n); }
minfft_aux* minfft_mkaux_t4_2d(int N1, int N2) { int n[2] = {N1, N2}; return mk(mfx::FAM_T4, 2, n); }
minfft_aux* minfft_mkaux_t4_3d(int N1, int N2, int N3) { int n[3] = {N1, N2, N3}; return mk(mfx::FAM_T4, 3, n); }

void minfft_free_aux(minfft_aux* a) {
    mfx::Plan* p = reinterpret_cast<mfx::Plan*>(a);
    if (!p || p->magic != 0x6d667870u) return;
    mfx::free_plan(p);
}

}  // extern "C"
