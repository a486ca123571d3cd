// engine.cpp -- plans, pass scheduling and launches for the transform engine.
//
// A d-dimensional transform is executed as one batched 1D pass per axis,
// last axis first, exactly the order in which the reference's mkcx/mkrx
// recursion applies its 1D routine (minfft.c:33-72); the real DFT runs its
// real pass along the last axis first and then complex passes over the
// leading axes (minfft.c:582-597), its inverse the other way round
// (minfft.c:662-687).  Each pass is one launch of the fused tile kernel when
// the axis length fits one work-group (2^14 FP32 / 2^13 FP64 points on a
// 160 KiB-LDS CU); at and just past that limit a radix-2-split single pass
// (split2.hpp) may take over; longer axes use a two-step (N = N1*N2)
// decomposition through a scratch buffer, or for batches of long 1D entries
// the fused two-step kernel (big.hpp).  Launch shapes come from the schedule
// in force (tune(): defaults, MINFFT_TUNE, or the autotuner's choice for the
// plan / batch group).
#include "engine.hpp"
#include "rlong.hpp"

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>

#include "bigtable.hpp"
#include "ktable.hpp"
#include "tile.hpp"
#include "big.hpp"
#include "coop.hpp"
#include "multi.hpp"
#include "autotune.hpp"

namespace mfx {

hipError_t launch_twiddles(double2* tw64, float2* tw32, long long total, hipStream_t st);
hipError_t launch_host_flag(unsigned* flag, unsigned v, hipStream_t st);

// ------------------------------------------------------------------ errors
static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }
const char* last_error() { return g_err.c_str(); }
void clear_error() { g_err.clear(); }

#define HIPCHK(expr, what)                                                           \
    do {                                                                             \
        hipError_t e_ = (expr);                                                      \
        if (e_ != hipSuccess) {                                                      \
            set_error(std::string(what) + ": " + hipGetErrorString(e_));             \
            return -1;                                                               \
        }                                                                            \
    } while (0)

int family_of(int kind) {
    switch (kind) {
        case K_DFT: case K_INVDFT: return FAM_DFT;
        case K_REALDFT: case K_INVREALDFT: return FAM_REALDFT;
        case K_DCT2: case K_DST2: case K_DCT3: case K_DST3: return FAM_T2T3;
        default: return FAM_T4;
    }
}

size_t real_size(Prec p) { return p == PREC_F64 ? 8 : 4; }

static int ilog2i(long long x) {
    int l = 0;
    while ((1LL << l) < x) ++l;
    return l;
}

// --------------------------------------------------------------- registry
#define DECL(k) void register_f32_k##k(KTable&); void register_f64_k##k(KTable&);
DECL(0) DECL(1) DECL(2) DECL(3) DECL(4) DECL(5) DECL(6) DECL(7) DECL(8) DECL(9)
#undef DECL

void register_f32(KTable& t) {
    register_f32_k0(t); register_f32_k1(t); register_f32_k2(t); register_f32_k3(t); register_f32_k4(t);
    register_f32_k5(t); register_f32_k6(t); register_f32_k7(t); register_f32_k8(t); register_f32_k9(t);
}
void register_f64(KTable& t) {
    register_f64_k0(t); register_f64_k1(t); register_f64_k2(t); register_f64_k3(t); register_f64_k4(t);
    register_f64_k5(t); register_f64_k6(t); register_f64_k7(t); register_f64_k8(t); register_f64_k9(t);
}

static BigTable g_big;
static std::once_flag g_big_once;
static const BigTable& bigtable() {
    std::call_once(g_big_once, [] {
        std::memset(&g_big, 0, sizeof(g_big));
        register_big(g_big);
        register_rsplit(g_big);
        register_csplit(g_big);
    });
    return g_big;
}

static KTable g_kt[2];
static std::once_flag g_kt_once;
static const KTable& ktable(Prec p) {
    std::call_once(g_kt_once, [] {
        std::memset(g_kt, 0, sizeof(g_kt));
        register_f64(g_kt[PREC_F64]);
        register_f32(g_kt[PREC_F32]);
    });
    return g_kt[p];
}

// ----------------------------------------------------------------- device
static Device g_dev;
static std::mutex g_dev_mu;

Device& device() { return g_dev; }

hipError_t dev_alloc(void** p, size_t bytes) {
    const char* e = std::getenv("MINFFT_FAIL_ALLOC");
    if (e && *e) {
        static std::mutex mu;
        static std::string seen;
        static long long count = 0;
        std::lock_guard<std::mutex> lk(mu);
        if (seen != e) {
            seen = e;
            count = 0;
        }
        const long long k = std::atoll(e);
        if (k > 0 && ++count >= k) {
            *p = nullptr;
            return hipErrorOutOfMemory;
        }
    }
    return hipMalloc(p, bytes);
}

bool ensure_device() {
    std::lock_guard<std::mutex> lk(g_dev_mu);
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess) {
        set_error("no HIP device available");
        return false;
    }
    if (g_dev.id == cur) return true;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, cur) != hipSuccess) {
        set_error("hipGetDeviceProperties failed");
        return false;
    }
    g_dev.id = cur;
    g_dev.ncu = prop.multiProcessorCount;
    int lds = (int)prop.sharedMemPerBlock;
    int optin = 0;
    if (hipDeviceGetAttribute(&optin, hipDeviceAttributeSharedMemPerBlockOptin, cur) == hipSuccess &&
        optin > lds)
        lds = optin;
    g_dev.lds_max = lds;
    g_dev.nxcd = 8;
    const KTable& k64 = ktable(PREC_F64);
    const KTable& k32 = ktable(PREC_F32);
    for (int p = 0; p < 2; ++p) {
        const size_t cs = p == PREC_F64 ? 16 : 8;
        const KTable& kt = p == PREC_F64 ? k64 : k32;
        int m = 0;
        for (int l = 0; l <= kt.maxlog; ++l)
            if ((size_t)slot_len(1 << l) * cs <= (size_t)lds) m = l;
        g_dev.maxlog[p] = m;
    }
    return true;
}

bool is_device_pointer(const void* p) {
    if (!p) return false;
    hipPointerAttribute_t a;
    hipError_t e = hipPointerGetAttributes(&a, p);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged ||
           a.type == hipMemoryTypeUnified;
}

// ------------------------------------------------------------ tile tuning
// Process-wide schedule: defaults, then MINFFT_TUNE="key=value,..." (keys:
// tpw, cfb, cfcore, wave, xcd, e32, e4, stage, chunk, split, split2, fused,
// wpc, falt, rsplit, coop, coopla, mall, rlong, mstages, csplit -- the fields
// of Tune in engine.hpp).
static Tune g_tune;
static std::once_flag g_tune_once;
static thread_local const Tune* t_tune = nullptr;

bool parse_tune(const char* spec, Tune& t) {
    std::string s(spec ? spec : "");
    size_t pos = 0;
    bool ok = true;
    while (pos < s.size()) {
        size_t c = s.find(',', pos);
        if (c == std::string::npos) c = s.size();
        std::string kv = s.substr(pos, c - pos);
        pos = c + 1;
        if (kv.empty()) continue;
        size_t eq = kv.find('=');
        if (eq == std::string::npos) {
            ok = false;
            continue;
        }
        std::string k = kv.substr(0, eq);
        int v = std::atoi(kv.c_str() + eq + 1);
        if (k == "tpw" && v > 0) t.tpw = v;
        else if (k == "cfb" && v > 0) t.cfb = v;
        else if (k == "wave") t.wave = v;
        else if (k == "xcd") t.xcd = v;
        else if (k == "stage") t.stage = v;
        else if (k == "chunk" && v > 0) t.chunk_mb = v;
        else if (k == "split") t.split = v;
        else if (k == "split2") t.split2 = v;
        else if (k == "cfcore") t.cfcore = v;
        else if (k == "e32") t.e32 = v;
        else if (k == "e4") t.e4 = v;
        else if (k == "fused") t.fused = v;
        else if (k == "wpc" && v > 0) t.fused_wpc = v;
        else if (k == "falt") t.fused_alt = v;
        else if (k == "rsplit") t.rsplit = v;
        else if (k == "coop") t.coop = v;
        else if (k == "coopla" && v >= 1 && v <= 64) t.coop_la = v;
        else if (k == "mall" && v >= 0) t.mall_mb = v;
        else if (k == "rlong") t.rlong = v;
        else if (k == "mstages" && (v == 1 || v == 2)) t.mstages = v;
        else if (k == "csplit") t.csplit = v;
        else ok = false;
    }
    return ok;
}

const Tune& base_tune() {
    std::call_once(g_tune_once, [] {
        const char* e = std::getenv("MINFFT_TUNE");
        if (e) (void)parse_tune(e, g_tune);
    });
    return g_tune;
}

const Tune& tune() { return t_tune ? *t_tune : base_tune(); }

TuneScope::TuneScope(const Tune* t) : prev(t_tune) { t_tune = t; }
TuneScope::~TuneScope() { t_tune = prev; }

// -------------------------------------------------------------- launching
static std::mutex g_attr_mu;
static std::map<std::pair<int, const void*>, size_t> g_attr_lds;  // (device, kernel) -> LDS allowed

// resident work-groups per CU of a kernel at this launch shape (cached)
static int occupancy(const void* fn, unsigned block, size_t lds) {
    static std::mutex mu;
    static std::map<std::tuple<int, const void*, unsigned, size_t>, int> cache;
    std::lock_guard<std::mutex> lk(mu);
    const auto key = std::make_tuple(device().id, fn, block, lds);
    const auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    if (lds > 65536) (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fn, (int)block, lds) != hipSuccess || n < 1) {
        (void)hipGetLastError();
        n = 1;
    }
    cache[key] = n;
    return n;
}

static int launch_fn(const void* fn, unsigned grid, unsigned block, void** args, size_t lds,
                     hipStream_t st) {
    if (lds > 65536) {
        // dynamic LDS beyond 64 KiB must be allowed per kernel (once per device)
        std::lock_guard<std::mutex> lk(g_attr_mu);
        size_t& have = g_attr_lds[{device().id, fn}];
        if (have < lds) {
            HIPCHK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds),
                   "hipFuncSetAttribute");
            have = lds;
        }
    }
    HIPCHK(hipLaunchKernel(fn, dim3(grid), dim3(block), args, lds, st), "kernel launch");
    return 0;
}

// Place a pass's batch levels into TileArgs.  TileArgs levels are
// right-aligned (level 3 fastest); an entry-offset table always belongs to
// TileArgs level 0, so a Levels whose slowest level carries a table keeps
// that level at 0 and right-aligns the rest.  *fs_level (an index into lv,
// or -1) is translated the same way.
static void fill_levels(TileArgs& A, const Levels& lv, int* fs_level = nullptr) {
    for (int k = 0; k < 4; ++k) {
        A.B[k] = 1;
        A.in_s[k] = A.out_s[k] = 0;
    }
    const bool tab = (lv.ioff || lv.ooff) && lv.n > 0;
    const int off = 4 - lv.n;
    for (int k = 0; k < lv.n; ++k) {
        const int dst = (tab && k == 0) ? 0 : off + k;
        A.B[dst] = lv.B[k];
        A.in_s[dst] = lv.is[k];
        A.out_s[dst] = lv.os[k];
    }
    if (fs_level && *fs_level >= 0) *fs_level = (tab && *fs_level == 0) ? 0 : off + *fs_level;
    A.in_off = tab ? lv.ioff : nullptr;
    A.out_off = tab ? lv.ooff : nullptr;
    for (int k = 0; k < 4; ++k) {
        const unsigned b = A.B[k];
        A.lgB[k] = (b & (b - 1)) == 0 ? ilog2i((int)b) : -1;
    }
    A.C = 1;
    A.lgC = 0;
}

// A request that the next one-work-group tile launch of this thread store a
// completion word for the host (signal_host in tile.hpp); `used` tells
// whether a launch took it.
struct HostSignal {
    unsigned* flag = nullptr;
    unsigned value = 0;
    bool used = false;
};
static thread_local HostSignal* t_signal = nullptr;

// Passes over at least this many bytes use non-temporal I/O where the
// kernel has the form (the MALL of MI355X holds 256 MiB)
constexpr unsigned long long kStreamBytes = 256ull << 20;

// One launch of the fused tile kernel.
static int launch_tile(Prec prec, int kind, int logL, const void* tw, const void* tw64, const void* in,
                       long long in_e,
                       void* out, long long out_e, const Levels& lv, int fs_level, int fs_logN,
                       hipStream_t st, bool split16 = false, bool pack = false) {
    const KTable& kt = ktable(prec);
    Device& D = device();
    if (logL > D.maxlog[prec]) {
        set_error("internal: tile length beyond the single-tile limit");
        return -1;
    }
    split16 = split16 && prec == PREC_F32 && logL == 9 && (kind == K_DFT || kind == K_INVDFT) && fs_logN == 0;
    const bool e32 = !split16 && tune().e32 && kt.tile32[kind][logL];
    // a small complex pass (a few transforms in all) is bound by latency,
    // not bandwidth: 4 points per lane give each transform four times the
    // lanes of the 16-point tile (one 1024-point FP64 transform: 3.5 against
    // 4.0 us per launch back to back)
    const bool e4 = !split16 && !e32 && tune().e4 && (kind == K_DFT || kind == K_INVDFT) && fs_logN == 0 &&
                    kt.tile4[kind == K_INVDFT][logL] && lv.count() << logL <= (1ull << 16);
    const void* fn = e32 ? kt.tile32[kind][logL] : (e4 ? kt.tile4[kind == K_INVDFT][logL] : kt.tile[kind][logL]);
    const int E = split16 ? 16 : (e32 ? 32 : (e4 ? 4 : kt.tile_E[logL]));
    const int L = 1 << logL;
    const int TPT = L / E;
    TileArgs A;
    std::memset(&A, 0, sizeof(A));
    A.in = in;
    A.out = out;
    A.tw = tw;
    A.tw64 = (const double2*)tw64;
    A.in_e = in_e;
    A.out_e = out_e;
    fill_levels(A, lv, &fs_level);
    unsigned long long ntr = lv.count();
    if (ntr == 0) return 0;
    if (ntr > 0xffffffffull) {
        set_error("batch too large");
        return -1;
    }
    A.ntrans = (unsigned)ntr;
    A.fs_level = fs_level < 0 ? 0 : fs_level;
    A.fs_logN = fs_logN;

    int f = -1;
    for (int k = 3; k >= 0; --k)
        if (A.B[k] > 1) { f = k; break; }
    const bool tab_f = (f == 0) && (A.in_off || A.out_off);
    const bool in_row = in_e == 1, out_row = out_e == 1;
    const bool in_cf = f >= 0 && !tab_f && A.in_s[f] == 1;
    const bool out_cf = f >= 0 && !tab_f && A.out_s[f] == 1;
    const bool cplx = (kind == K_DFT || kind == K_INVDFT);
    const Tune& tn = tune();
    // core mapping of column passes: c-fast (complex kinds: direct I/O) or
    // row-mapped with the column I/O staged through LDS
    const bool cf_core = tn.cfcore < 0 ? cplx : tn.cfcore != 0;
    if (cplx && tn.stage) {
        A.in_mode = in_row ? IO_STAGE_ROW : IO_STAGE_CF;
        A.out_mode = out_row ? IO_STAGE_ROW : IO_STAGE_CF;
        A.cfast = (A.in_mode == IO_STAGE_CF && cf_core) ? 1 : 0;
    } else if (cplx && cf_core) {
        if (in_row && out_row) { A.cfast = 0; A.in_mode = IO_DIRECT; A.out_mode = IO_DIRECT; }
        else if (in_cf && out_cf) { A.cfast = 1; A.in_mode = IO_DIRECT; A.out_mode = IO_DIRECT; }
        else if (out_cf) { A.cfast = 1; A.out_mode = IO_DIRECT; A.in_mode = in_row ? IO_STAGE_ROW : IO_DIRECT; }
        else if (out_row) { A.cfast = 0; A.out_mode = IO_DIRECT; A.in_mode = in_cf ? IO_STAGE_CF : IO_DIRECT; }
        else { A.cfast = in_cf ? 1 : 0; A.in_mode = IO_DIRECT; A.out_mode = IO_DIRECT; }
    } else if (cplx) {
        A.cfast = 0;
        A.in_mode = in_row ? IO_DIRECT : (in_cf ? IO_STAGE_CF : IO_DIRECT);
        A.out_mode = out_row ? IO_DIRECT : (out_cf ? IO_STAGE_CF : IO_DIRECT);
    } else {
        A.in_mode = in_row ? IO_STAGE_ROW : (in_cf ? IO_STAGE_CF : IO_STAGE_ROW);
        A.out_mode = out_row ? IO_STAGE_ROW : (out_cf ? IO_STAGE_CF : IO_STAGE_ROW);
        A.cfast = (A.in_mode == IO_STAGE_CF && cf_core) ? 1 : 0;
    }
    // split-FP16 MFMA tiles: the column variant when both sides are columns
    // of adjacent transforms (direct I/O, 16 transforms per work-group) and
    // the schedule wants a c-fast core; otherwise the row variant (two
    // transforms per wavefront, direct I/O for contiguous rows, other I/O
    // staged through LDS)
    const bool s16cf = split16 && in_cf && out_cf && cf_core;
    if (s16cf) {
        fn = bigtable().mfma512cf[kind == K_INVDFT];
        A.cfast = 1;
        A.in_mode = IO_DIRECT;
        A.out_mode = IO_DIRECT;
    } else if (split16) {
        fn = bigtable().mfma512[kind == K_INVDFT];
        A.cfast = 0;
        A.in_mode = in_row ? IO_DIRECT : (in_cf ? IO_STAGE_CF : IO_STAGE_ROW);
        A.out_mode = out_row ? IO_DIRECT : (out_cf ? IO_STAGE_CF : IO_STAGE_ROW);
    }
    // packed pair (columns h = 0 and h = H-1 of a real DFT's half spectrum
    // as one complex column, tile.hpp): the c-fast direct form only; the
    // pass then covers h = 0..H-2
    if (pack && (kind == K_DFT || kind == K_INVDFT) && !split16 && !e32 && !e4 && fs_logN == 0 &&
        kt.tilepk[kind == K_INVDFT][logL] && A.cfast && A.in_mode == IO_DIRECT && A.out_mode == IO_DIRECT &&
        f == 3 && A.B[3] >= 2 && A.in_s[3] == 1 && A.out_s[3] == 1) {
        fn = kt.tilepk[kind == K_INVDFT][logL];
        A.pack_in = (long long)(A.B[3] - 1);
        A.pack_out = (long long)(A.B[3] - 1);
        ntr = ntr / A.B[3] * (A.B[3] - 1);
        A.B[3] -= 1;
        A.lgB[3] = (A.B[3] & (A.B[3] - 1)) == 0 ? ilog2i((int)A.B[3]) : -1;
        A.ntrans = (unsigned)ntr;
    }
    const size_t cs = prec == PREC_F64 ? 16 : 8;
    // radix-2-split column tile (csplit.hpp): both sides are columns of
    // adjacent transforms with direct c-fast I/O; half of each transform in
    // the LDS at a time, so a work-group takes twice the columns of the
    // whole-transform tile (twice the row segment) in the same LDS.  By
    // default for FP64 column passes and for FP32 512-point ones with
    // strides of 1 MiB and more; 512^3 per pass (kernel trace): FP32 outer
    // axis (2 MiB) 473 -> 433 us, middle axis (4 KiB) 370 -> 390 us (so kept
    // off there); FP64 outer 987 -> 847 us, middle 750 -> 723 us
    {
        const BigTable& bt = bigtable();
        const int pi = prec == PREC_F64 ? 0 : 1, iv = kind == K_INVDFT;
        const long long far_e = (long long)(1 << 20) / (long long)cs;
        // (FP32 1024-point columns keep the whole-transform tile by default:
        // its 16 columns already span 128 bytes; 1024^3 outer axis 11.67-11.80
        // against 11.82-11.89 ms per transform with the split tile)
        const bool cs_on = tn.csplit > 0 ||
                           (tn.csplit < 0 && (prec == PREC_F64 || (logL == 9 && (in_e >= far_e || out_e >= far_e))));
        // split-FP16: the matrix-core form of the same tile (mfma.hpp), 32
        // transforms per work-group, in place of the 16-transform column variant
        if (cs_on && s16cf && !in_row && !out_row && tn.mstages != 2 && bt.mfma512cs[0][iv]) {
            const int C = 32;
            const bool nt = ntr * (unsigned long long)L * cs >= kStreamBytes;
            A.C = C;
            A.lgC = 5;
            A.sc = slot_len(L / 2);
            const unsigned long long tiles = (ntr + C - 1) / C;
            A.ntiles = (unsigned)tiles;
            const bool far_stride = in_e >= far_e || out_e >= far_e;
            A.xcd = (tn.xcd && !far_stride && tiles >= 2ull * D.nxcd) ? D.nxcd : 1;
            A.cfast = 1;
            A.wave = 0;
            void* args[] = {&A};
            return launch_fn(bt.mfma512cs[nt ? 1 : 0][iv], (unsigned)tiles, 512u, args, (size_t)C * A.sc * cs, st);
        }
        if (cs_on && cplx && !split16 && !e32 && !e4 && fs_logN == 0 && fn == kt.tile[kind][logL] &&
            A.cfast && A.in_mode == IO_DIRECT && A.out_mode == IO_DIRECT && in_cf && out_cf && !in_row &&
            !out_row && bt.csplit[pi][logL][0][iv]) {
            const int tpt = bt.csplit_tpt[pi][logL];
            const size_t slot = bt.csplit_slot[pi][logL];
            int C = std::max(1, (int)(2 * (size_t)tn.cfb / cs));
            while (C > 1 && C * tpt > 512) C >>= 1;  // the kernel's launch bound
            if ((unsigned long long)C > ntr) C = (int)ntr;
            while (C > 1 && (C / 2) * tpt >= 64 && (ntr + C - 1) / C < 2ull * (unsigned long long)D.ncu) C >>= 1;
            while (C > 1 && (size_t)C * slot > (size_t)D.lds_max) C >>= 1;
            while (C & (C - 1)) C &= C - 1;
            const bool nt = ntr * (unsigned long long)L * cs >= kStreamBytes && (size_t)C * cs >= 128;
            A.C = C;
            A.lgC = ilog2i(C);
            A.sc = slot_len(L / 2);
            const unsigned long long tiles = (ntr + C - 1) / C;
            A.ntiles = (unsigned)tiles;
            const bool far_stride = in_e >= far_e || out_e >= far_e;
            A.xcd = (tn.xcd && !far_stride && tiles >= 2ull * D.nxcd) ? D.nxcd : 1;
            A.wave = 0;
            if (t_signal && !t_signal->used && tiles == 1) {
                A.hflag = t_signal->flag;
                A.hval = t_signal->value;
                t_signal->used = true;
            }
            void* args[] = {&A};
            return launch_fn(bt.csplit[pi][logL][nt ? 1 : 0][iv], (unsigned)tiles, (unsigned)(C * tpt), args,
                             (size_t)C * slot, st);
        }
    }
    const bool wants_cf = A.cfast || A.in_mode == IO_STAGE_CF || A.out_mode == IO_STAGE_CF;
    const int sc = slot_len(L);
    const int maxT = (prec == PREC_F64 || e32 || split16) ? 512 : 1024;  // the kernels' launch bounds
    int C;
    if (wants_cf) {
        // several columns per tile so that each row segment of the strided
        // side spans cfb bytes (real items for the DCT/DST kinds)
        const size_t isz = cplx || kind == K_REALDFT || kind == K_INVREALDFT ? cs : cs / 2;
        C = std::max(1, (int)(tn.cfb / isz));
        while (C > 1 && C * TPT > maxT) C >>= 1;
        // staged (real-kind) column tiles: two work-groups per CU rather than
        // one twice as wide, while the row segment keeps 64 bytes (256 images
        // of 1024^2 FP32 DCT-2/DST-2/DCT-4: 0.964 -> 0.948 ms per batch, the
        // DCT-2 column pass alone 185 -> 174 us; the tuner's subset timing
        // did not resolve the difference)
        if (!cplx)
            while (C > 1 && (size_t)C * sc * cs > (size_t)D.lds_max / 2 && (size_t)(C / 2) * isz >= 64) C >>= 1;
    } else {
        C = std::max(1, std::min(tn.tpw, maxT) / TPT);
    }
    if ((unsigned long long)C > ntr) C = (int)ntr;
    // small batches: fewer transforms per tile until every CU gets two
    // work-groups (wider row segments are worth less than idle CUs), down to
    // one wavefront per work-group
    while (C > 1 && (C / 2) * TPT >= 64 && (ntr + C - 1) / C < 2ull * (unsigned long long)D.ncu) C >>= 1;
    size_t budget = (size_t)D.lds_max;
    while (C > 1 && (size_t)C * sc * cs > budget) C >>= 1;
    if ((size_t)C * sc * cs > budget) {
        set_error("internal: LDS budget exceeded");
        return -1;
    }
    // a power of two (the kernels split the lane index with shifts)
    while (C & (C - 1)) C &= C - 1;
    if (split16 && C < 2) C = 2;  // MFMA operands span a full wavefront: two transforms
    if (s16cf) C = 16;            // the column variant's layout: 16 transforms, 512 lanes
    // A complex pass over more data than the MALL holds streams through it
    // anyway: non-temporal direct I/O keeps it from displacing what is
    // re-read (512^3 FP32: 1.30 -> 1.24 ms).  Smaller passes keep the MALL's
    // reuse between passes (2048^2 FP64 passes were 12-24 % slower
    // non-temporal), and so do column tiles whose row segments are shorter
    // than a 128-byte line: a non-temporal access does not keep the rest of
    // the line for the neighbouring tile (32-byte segments: 4x slower).
    const bool stream = ntr * (unsigned long long)L * cs >= kStreamBytes && (!wants_cf || (size_t)C * cs >= 128);
    if (stream) {
        const BigTable& bt = bigtable();
        if (fn == bt.mfma512cf[kind == K_INVDFT]) fn = bt.mfma512cfnt[kind == K_INVDFT];
        else if (fn == bt.mfma512[kind == K_INVDFT]) fn = bt.mfma512nt[kind == K_INVDFT];
        else if ((kind == K_DFT || kind == K_INVDFT) && fn == kt.tile[kind][logL] && kt.tilent[kind == K_INVDFT][logL])
            fn = kt.tilent[kind == K_INVDFT][logL];
    }
    A.C = C;
    A.lgC = ilog2i(C);
    A.sc = sc;
    const unsigned long long tiles = (ntr + C - 1) / C;
    A.ntiles = (unsigned)tiles;
    // XCD-grouped column tiles share their rows' lines in one L2; for
    // strides of 1 MiB and more (the outer axis of a 512^3 array: 2 MiB)
    // the grouping measured slower -- 512^3 split-FP16 outer pass 537 ->
    // 516 us without it, the FP32 tile 475 -> 473 us, while the middle axis
    // (4 KiB) keeps it (FP32 370 against 391 us without)
    const long long far = (long long)(1 << 20) / (long long)cs;
    const bool far_stride = (!in_row && A.in_e >= far) || (!out_row && A.out_e >= far);
    A.xcd = (tn.xcd && wants_cf && !far_stride && tiles >= 2ull * D.nxcd) ? D.nxcd : 1;
    // wavefront-local core exchanges need a row-mapped core; staged column
    // I/O keeps its work-group barriers (tile.hpp)
    A.wave = (tn.wave && !A.cfast) ? 1 : 0;
    const BigTable& btab = bigtable();
    // both radix-16 stages on the matrix cores (schedule key mstages=2)
    if (split16 && tn.mstages == 2) {
        const int iv = kind == K_INVDFT;
        for (int cf = 0; cf < 2; ++cf)
            for (int nt = 0; nt < 2; ++nt) {
                const void* one = cf ? (nt ? btab.mfma512cfnt[iv] : btab.mfma512cf[iv])
                                     : (nt ? btab.mfma512nt[iv] : btab.mfma512[iv]);
                if (fn == one) fn = btab.mfma2[cf][nt][iv];
            }
    }
    const bool plain_tile = !split16;
    if (t_signal && !t_signal->used && tiles == 1 && plain_tile) {
        A.hflag = t_signal->flag;
        A.hval = t_signal->value;
        t_signal->used = true;
    }
    void* args[] = {&A};
    return launch_fn(fn, (unsigned)tiles, (unsigned)(C * TPT), args, (size_t)C * sc * cs, st);
}

static int launch_elementwise(const void* fn, TileArgs& A, int L, long long total, hipStream_t st) {
    if (total <= 0) return 0;
    const unsigned bs = 256;
    const unsigned long long grid = ((unsigned long long)total + bs - 1) / bs;
    void* args[] = {&A, &L, &total};
    return launch_fn(fn, (unsigned)grid, bs, args, 0, st);
}

int needed_tw_log(int kind, int logN) {
    if (kind == K_DFT || kind == K_INVDFT) return std::max(logN, 1);
    return std::max(logN + 3, 1);
}

const char* axis_length_error(Prec prec, int kind, int logN) {
    if (!ensure_device()) return "no GPU device";
    const int maxlog = device().maxlog[prec];
    // the complex FFT the pass runs: the axis itself, or half of it for the
    // real, DCT and DST kinds
    const int lc = (kind == K_DFT || kind == K_INVDFT) ? logN : logN - 1;
    if (lc > maxlog && lc - lc / 2 > maxlog) return "transform length exceeds the two-step limit";
    return nullptr;
}

size_t axis_scratch_bytes(Prec prec, int kind, int logN, unsigned long long ntrans,
                          unsigned long long in_extent) {
    if (!ensure_device()) return 0;
    const int maxlog = device().maxlog[prec];
    const size_t cs = prec == PREC_F64 ? 16 : 8;
    if (kind == K_DFT || kind == K_INVDFT) {
        if (logN <= maxlog) return 0;
        return (size_t)in_extent * cs;
    }
    if (logN == 0 || logN - 1 <= maxlog) return 0;
    return (size_t)ntrans * ((size_t)1 << (logN - 1)) * cs;
}

// complex pass along an axis of element stride e (same on both sides)
// single-pass radix-2-split launch (split2.hpp) for contiguous transforms of
// length 2^logL, if instantiated and its LDS fits; 1 = not applicable
static int launch_split2(const Workspace& ws, Prec prec, int kind, int logL, const void* in, long long in_e,
                         void* out, long long out_e, const Levels& lv, hipStream_t st) {
    const BigTable& bt = bigtable();
    const int pi = prec == PREC_F64 ? 0 : 1;
    if (in_e != 1 || out_e != 1 || logL > 16 || !bt.split2[pi][logL][0][0] ||
        bt.split2_lds[pi][logL] > (size_t)device().lds_max)
        return 1;
    const int v = (tune().e32 && bt.split2[pi][logL][1][0]) ? 1 : 0;
    TileArgs A;
    std::memset(&A, 0, sizeof(A));
    A.in = in;
    A.out = out;
    A.tw = ws.tw;
    A.tw64 = (const double2*)ws.tw64;
    A.in_e = 1;
    A.out_e = 1;
    fill_levels(A, lv);
    const unsigned long long ntr = lv.count();
    if (ntr == 0) return 0;
    if (ntr > 0x7fffffffull) {
        set_error("batch too large");
        return -1;
    }
    A.ntrans = (unsigned)ntr;
    A.ntiles = (unsigned)ntr;
    A.C = 1;
    A.xcd = 1;
    // non-temporal I/O for whole 1D transforms and passes larger than the MALL
    const bool nt = lv.stream_io || ntr * (1ull << logL) * (prec == PREC_F64 ? 16u : 8u) >= kStreamBytes;
    void* args[] = {&A};
    return launch_fn((nt ? bt.split2nt : bt.split2)[pi][logL][v][kind == K_INVDFT], (unsigned)ntr,
                     (unsigned)bt.split2_bt[pi][logL][v],
                     args, bt.split2_lds[pi][logL], st);
}

// single-pass residue-split launch (rsplit.hpp) for contiguous transforms of
// length 2^logL, with the schedule's split (tune().rsplit = 10*log2 P + B),
// if instantiated and its LDS fits; 1 = not applicable
static int launch_rsplit(const Workspace& ws, Prec prec, int kind, int logL, const void* in, long long in_e,
                         void* out, long long out_e, const Levels& lv, hipStream_t st) {
    const int code = tune().rsplit;
    const int lp = code / 10, b = code % 10;
    if (code <= 0 || lp < 1 || lp > 4 || b < 1 || b > 2 || in_e != 1 || out_e != 1 || logL > 16) return 1;
    // the P/B work-groups of a transform all read its whole input while
    // they write their outputs: never in place
    if (in == out) return 1;
    const BigTable& bt = bigtable();
    const int pi = prec == PREC_F64 ? 0 : 1;
    const void* fn = bt.rsplit[pi][logL][lp - 1][b - 1][kind == K_INVDFT];
    const size_t lds = bt.rsplit_lds[pi][logL][lp - 1][b - 1];
    if (!fn || lds > (size_t)device().lds_max || ws.tw_log < logL) return 1;
    TileArgs A;
    std::memset(&A, 0, sizeof(A));
    A.in = in;
    A.out = out;
    A.tw = ws.tw;
    A.tw64 = (const double2*)ws.tw64;
    A.in_e = 1;
    A.out_e = 1;
    fill_levels(A, lv);
    const unsigned long long ntr = lv.count();
    if (ntr == 0) return 0;
    // P/B work-groups per transform, transforms in groups of 8 (one per XCD)
    const unsigned long long G = (1ull << lp) / (unsigned)b;
    const unsigned long long grid = ((ntr + 7) / 8) * 8 * G;
    if (grid > 0x7fffffffull) {
        set_error("batch too large");
        return -1;
    }
    A.ntrans = (unsigned)ntr;
    A.ntiles = (unsigned)grid;
    A.C = 1;
    A.xcd = 1;
    void* args[] = {&A};
    return launch_fn(fn, (unsigned)grid, (unsigned)bt.rsplit_bt[pi][logL][lp - 1][b - 1], args, lds, st);
}

// Queues of the XCD-cooperative kernel: one per XCD, which must each have
// an L2 of their own -- 8 on a whole MI355X (256 CUs, 32 per XCD), 1 on a
// one-XCD partition; 0 (kernel not used) for any other layout.
unsigned coop_queues() {
    if (!ensure_device()) return 0;
    const int ncu = device().ncu;
    if (ncu == 256) return 8;
    if (ncu == 32) return 1;
    return 0;
}

int multi_grid(Prec prec) {
    if (!ensure_device()) return 0;
    const BigTable& bt = bigtable();
    const int pi = prec == PREC_F64 ? 0 : 1;
    if (!bt.multi[pi][0] || bt.multi_lds[pi] > (size_t)device().lds_max) return 0;
    // resident work-groups of both directions (a larger grid would still be
    // correct: late work-groups find the items taken)
    const int occ = std::min(occupancy(bt.multi[pi][0], kMultiThreads, bt.multi_lds[pi]),
                             occupancy(bt.multi[pi][1], kMultiThreads, bt.multi_lds[pi]));
    return occ * device().ncu;
}

int launch_multi(Prec prec, bool inv, const MultiArgs& a, unsigned grid, bool per_item, hipStream_t st) {
    const BigTable& bt = bigtable();
    const int pi = prec == PREC_F64 ? 0 : 1;
    MultiArgs m = a;
    if (per_item) {
        // one work-group per item: the grid is the item count
        m.counter = nullptr;
        void* args[] = {&m};
        return launch_fn(bt.multi[pi][inv ? 1 : 0], a.total, kMultiThreads, args, bt.multi_lds[pi], st);
    }
    unsigned* ctr = a.counter;
    void* rargs[] = {&ctr};
    if (launch_fn(bt.multi_reset, 1, 1, rargs, 0, st)) return -1;
    void* args[] = {&m};
    return launch_fn(bt.multi[pi][inv ? 1 : 0], grid, kMultiThreads, args, bt.multi_lds[pi], st);
}

std::vector<int> rsplit_codes(Prec prec, int logN) {
    std::vector<int> v;
    if (logN < 1 || logN > 16 || !ensure_device()) return v;
    const BigTable& bt = bigtable();
    const int pi = prec == PREC_F64 ? 0 : 1;
    for (int lp = 1; lp <= 4; ++lp)
        for (int b = 1; b <= 2; ++b)
            if (bt.rsplit[pi][logN][lp - 1][b - 1][0] && bt.rsplit_lds[pi][logN][lp - 1][b - 1] <= (size_t)device().lds_max)
                v.push_back(10 * lp + b);
    return v;
}

bool c2c_single_pass(Prec prec, int logL, const Tune& t, bool in_place) {
    if (!ensure_device()) return false;
    const int maxlog = device().maxlog[prec];
    if (logL <= maxlog) return true;
    if (logL > 16) return false;
    const BigTable& bt = bigtable();
    const int pi = prec == PREC_F64 ? 0 : 1;
    const size_t lmax = (size_t)device().lds_max;
    const int lp = t.rsplit / 10, b = t.rsplit % 10;
    if (!in_place && t.rsplit > 0 && lp >= 1 && lp <= 4 && b >= 1 && b <= 2 &&
        bt.rsplit[pi][logL][lp - 1][b - 1][0] && bt.rsplit_lds[pi][logL][lp - 1][b - 1] <= lmax)
        return true;
    const int sp2 = t.split2 >= 0 ? t.split2 : (prec == PREC_F64 ? 2 : 3);
    return sp2 >= 1 && logL == maxlog + 1 && bt.split2[pi][logL][0][0] && bt.split2_lds[pi][logL] <= lmax;
}

static int c2c_axis(const Workspace& ws, Prec prec, int kind, int logL, const void* in, long long in_e,
                    void* out, long long out_e, const Levels& lv, hipStream_t st) {
    const int maxlog = device().maxlog[prec];
    if (tune().rsplit > 0) {
        const int rc = launch_rsplit(ws, prec, kind, logL, in, in_e, out, out_e, lv, st);
        if (rc <= 0) return rc;
    }
    // radix-2-split kernel: always one past the tile limit (split2 >= 1),
    // at the limit with split2 >= 2, one below it with split2 >= 3
    const int sp2 = tune().split2 >= 0 ? tune().split2 : (prec == PREC_F64 ? 2 : 3);
    if ((sp2 >= 1 && logL == maxlog + 1) || (sp2 >= 2 && logL == maxlog) || (sp2 >= 3 && logL == maxlog - 1)) {
        const int rc = launch_split2(ws, prec, kind, logL, in, in_e, out, out_e, lv, st);
        if (rc <= 0) return rc;
    }
    if (logL <= maxlog)
        return launch_tile(prec, kind, logL, ws.tw, ws.tw64, in, in_e, out, out_e, lv, -1, 0, st, ws.split16 != 0,
                           lv.pack);
    if (in_e != out_e) {
        set_error("internal: two-step pass needs equal strides");
        return -1;
    }
    const long long e = in_e;
    // N = N1 * N2; the schedule may move the split off the middle
    const int l1 = std::max(std::max(1, logL - maxlog), std::min(logL / 2 + tune().split, std::min(maxlog, logL - 1)));
    const int l2 = logL - l1;
    if (l2 > maxlog) {
        set_error("transform length exceeds the two-step limit");
        return -1;
    }
    if (!ws.fbuf || !ws.tw) {
        set_error("internal: missing two-step scratch");
        return -1;
    }
    const long long N1 = 1LL << l1, N2 = 1LL << l2;
    if (lv.n >= 4) {
        set_error("internal: too many batch levels");
        return -1;
    }
    // step 1: length-N1 columns (stride N2*e), scratch mirrors the input
    // index space; twiddle W_N^{n2*k1}
    Levels s1, s2;
    const long long* fo = lv.foff ? lv.foff : lv.ioff;
    s1.ioff = lv.ioff;
    s1.ooff = fo;
    s2.ioff = fo;
    s2.ooff = lv.ooff;
    int fs_pos;
    if (e == 1) {
        for (int k = 0; k < lv.n; ++k) {
            s1.push(lv.B[k], lv.is[k], lv.is[k]);
            s2.push(lv.B[k], lv.is[k], lv.os[k]);
        }
        fs_pos = s1.n;
        s1.push((unsigned)N2, 1, 1);
        s2.push((unsigned)N1, N2, 1);
    } else {
        for (int k = 0; k + 1 < lv.n; ++k) {
            s1.push(lv.B[k], lv.is[k], lv.is[k]);
            s2.push(lv.B[k], lv.is[k], lv.os[k]);
        }
        fs_pos = s1.n;
        s1.push((unsigned)N2, e, e);
        s2.push((unsigned)N1, N2 * e, e);
        if (lv.n >= 1) {
            s1.push(lv.B[lv.n - 1], lv.is[lv.n - 1], lv.is[lv.n - 1]);
            s2.push(lv.B[lv.n - 1], lv.is[lv.n - 1], lv.os[lv.n - 1]);
        }
    }
    if ((lv.ioff || lv.ooff) && fs_pos == 0) {
        set_error("internal: offset table level clash");
        return -1;
    }
    // Batches of contiguous 1D entries with a slot per entry in the scratch:
    // the XCD-cooperative two-step kernel (coop.hpp), intermediates in L2.
    {
        const BigTable& bt = bigtable();
        const Tune& tn = tune();
        const int pi = prec == PREC_F64 ? 0 : 1;
        const unsigned cnt = lv.n ? lv.B[0] : 1;
        const unsigned nq = coop_queues();
        const void* fn = logL <= 16 ? bt.coop[pi][logL][kind == K_INVDFT] : nullptr;
        const size_t cs = prec == PREC_F64 ? 16 : 8;
        const bool coop_on = tn.coop < 0 ? prec == PREC_F64 : tn.coop != 0;
        if (fn && coop_on && nq && lv.n == 1 && lv.ioff && lv.ooff && lv.foff && e == 1 && ws.cbuf &&
            cnt >= (unsigned)device().ncu && cnt < (1u << 24) &&
            (size_t)(nq + cnt) * sizeof(unsigned) <= ws.cbuf_bytes &&
            (size_t)cnt * ((size_t)1 << logL) * cs <= ws.fs_bytes &&
            bt.coop_lds[pi][logL] <= (size_t)device().lds_max) {
            unsigned* ctr = ws.cbuf;
            unsigned nctr = nq + cnt;
            void* rargs[] = {&ctr, &nctr};
            if (launch_fn(bt.coop_reset, (nctr + 255) / 256, 256, rargs, 0, st)) return -1;
            CoopArgs C;
            C.in = in;
            C.out = out;
            C.scratch = ws.fbuf;
            C.tw = ws.tw;
            C.tw64 = (const double2*)ws.tw64;
            C.in_off = lv.ioff;
            C.out_off = lv.ooff;
            C.slot_off = lv.foff;
            C.ctr = ctr;
            C.count = cnt;
            C.nq = nq;
            C.lookahead = (unsigned)tn.coop_la;
            const unsigned grid = (unsigned)(device().ncu * std::max(1, bt.coop_wpc[pi][logL]));
            void* args[] = {&C};
            return launch_fn(fn, grid, (unsigned)bt.coop_bt[pi][logL], args, bt.coop_lds[pi][logL], st);
        }
    }
    // Batches of contiguous 1D entries: one fused launch, a work-group per
    // transform, intermediate in a per-work-group scratch slot.
    {
        const BigTable& bt = bigtable();
        const Tune& tn = tune();
        const int pi = prec == PREC_F64 ? 0 : 1;
        const int lg = std::min(logL, 16);
        const bool alt = tn.fused_alt && bt.alt[pi][lg][0];
        const void* fn = (logL <= 16) ? (alt ? bt.alt[pi][lg][kind == K_INVDFT] : bt.fn[pi][lg][kind == K_INVDFT])
                                      : nullptr;
        const int bwpc = alt ? bt.alt_wpc[pi][lg] : bt.wpc[pi][lg];
        const unsigned bthreads = (unsigned)(alt ? bt.alt_bt[pi][lg] : bt.bt[pi][lg]);
        const size_t blds = alt ? bt.alt_lds[pi][lg] : bt.lds[pi][lg];
        const unsigned cnt = lv.n ? lv.B[0] : 1;
        const int wpc = std::max(1, std::min(tn.fused_wpc, bwpc));
        const unsigned grid = std::min<unsigned>(cnt, (unsigned)(device().ncu * wpc));
        const size_t need = (size_t)grid * ((size_t)1 << logL) * (prec == PREC_F64 ? 16 : 8);
        if (fn && tn.fused && lv.n == 1 && lv.ioff && lv.ooff && e == 1 && cnt >= (unsigned)device().ncu &&
            need <= ws.fs_bytes) {
            BigArgs B;
            B.in = in;
            B.out = out;
            B.scratch = ws.fbuf;
            B.tw = ws.tw;
            B.tw64 = (const double2*)ws.tw64;
            B.in_off = lv.ioff;
            B.out_off = lv.ooff;
            B.count = cnt;
            void* args[] = {&B};
            return launch_fn(fn, grid, bthreads, args, blds, st);
        }
    }
    // With an entry table, run the two steps chunk by chunk (~64 MiB of
    // data per chunk) so each chunk's intermediate is still in the MALL
    // (infinity cache) when step 2 reads it back.
    const size_t cs = prec == PREC_F64 ? 16 : 8;
    const unsigned long long per_entry = lv.count() / (lv.n ? lv.B[0] : 1);
    unsigned chunk = lv.n ? lv.B[0] : 1;
    if ((lv.ioff || lv.ooff) && lv.n > 0) {
        const size_t target = (size_t)tune().chunk_mb << 20;
        const size_t ebytes = (size_t)per_entry * ((size_t)1 << logL) * cs;
        chunk = (unsigned)std::max<size_t>(1, target / std::max<size_t>(ebytes, 1));
        chunk = std::min(chunk, lv.B[0]);
    }
    const unsigned total = lv.n ? lv.B[0] : 1;
    for (unsigned c0 = 0; c0 < total; c0 += chunk) {
        const unsigned cn = std::min(chunk, total - c0);
        Levels a = s1, b = s2;
        if (lv.n && (lv.ioff || lv.ooff)) {
            a.B[0] = b.B[0] = cn;
            // scratch offsets: with a compact table every chunk reuses the
            // same (first) scratch entries, keeping the footprint one chunk
            const long long* fsc = lv.foff ? fo : (fo ? fo + c0 : nullptr);
            a.ioff = s1.ioff ? s1.ioff + c0 : nullptr;
            a.ooff = fsc;
            b.ioff = fsc;
            b.ooff = s2.ooff ? s2.ooff + c0 : nullptr;
        } else if (c0 > 0) {
            break;
        }
        int rc = launch_tile(prec, kind, l1, ws.tw, ws.tw64, in, N2 * e, ws.fbuf, N2 * e, a, fs_pos, logL, st);
        if (rc) return rc;
        rc = launch_tile(prec, kind, l2, ws.tw, ws.tw64, ws.fbuf, e, out, N1 * e, b, -1, 0, st);
        if (rc) return rc;
    }
    return 0;
}

// A real/DCT/DST axis whose half-length FFT (L = 2^logL) is longer than a
// tile, in two passes (rlong.hpp): L = N1*N2 with log2 N1 = logL/2; step 1
// reads the input through the pass's levels (or entry table) into the
// scratch ws.gbuf (transform b at b*L), step 2 writes the output through
// them.  Returns 1 without launching for lengths outside the instantiated
// step kernels.
static int rlong_axis(const Workspace& ws, Prec prec, int kind, int logL, const void* in, long long in_e,
                      void* out, long long out_e, const Levels& lv, hipStream_t st) {
    const int l1 = logL / 2, l2 = logL - l1;
    if (l1 < kRlongLogMin || l2 < kRlongLogMin || l1 > kRlongLogMax || l2 > kRlongLogMax) return 1;
    const KTable& kt = ktable(prec);
    const void* f1 = kt.rstep1[kind][l1];
    const void* f2 = kt.rstep2[kind][l2];
    if (!f1 || !f2) return 1;
    const Device& D = device();
    const size_t cs = prec == PREC_F64 ? 16 : 8;
    const unsigned long long ntr = lv.count();
    // columns (step 1) / rows (step 2) per work-group: up to 512 lanes, LDS
    // within the device limit, at least two rows (one pair) in step 2
    auto per_tile = [&](int lfft, int lother, int cmin) {
        const int tpt = (1 << lfft) / default_E(lfft);
        const size_t slot = (size_t)slot_len(1 << lfft) * cs;
        int C = std::min(1 << lother, std::max(1, 512 / tpt));
        while (C > cmin && (size_t)C * slot > (size_t)D.lds_max) C >>= 1;
        return (C >= cmin && (size_t)C * slot <= (size_t)D.lds_max) ? C : 0;
    };
    const int C1 = per_tile(l1, l2, 1), C2 = per_tile(l2, l1, 2);
    if (!C1 || !C2) return 1;
    const unsigned long long t1 = ntr << (l2 - ilog2i(C1)), t2 = ntr << (l1 - ilog2i(C2));
    if (t1 > 0xffffffffull || t2 > 0xffffffffull) return 1;

    TileArgs A;
    std::memset(&A, 0, sizeof(A));
    fill_levels(A, lv);
    A.ntrans = (unsigned)ntr;
    A.tw = ws.tw;
    A.tw64 = (const double2*)ws.tw64;
    A.fs_logN = logL;
    A.in = in;
    A.in_e = in_e;
    A.out = ws.gbuf;
    A.out_off = nullptr;
    A.C = C1;
    A.lgC = ilog2i(C1);
    A.sc = slot_len(1 << l1);
    A.ntiles = (unsigned)t1;
    A.xcd = t1 >= 2ull * D.nxcd ? D.nxcd : 1;
    const unsigned b1 = (unsigned)(C1 * ((1 << l1) / default_E(l1)));
    void* a1[] = {&A};
    int rc = launch_fn(f1, (unsigned)t1, b1, a1, (size_t)C1 * A.sc * cs, st);
    if (rc) return rc;

    TileArgs B;
    std::memset(&B, 0, sizeof(B));
    fill_levels(B, lv);
    B.ntrans = (unsigned)ntr;
    B.tw = ws.tw;
    B.tw64 = (const double2*)ws.tw64;
    B.fs_logN = logL;
    B.in = ws.gbuf;
    B.in_off = nullptr;
    B.out = out;
    B.out_e = out_e;
    B.C = C2;
    B.lgC = ilog2i(C2);
    B.sc = slot_len(1 << l2);
    B.ntiles = (unsigned)t2;
    B.xcd = t2 >= 2ull * D.nxcd ? D.nxcd : 1;
    const unsigned b2 = (unsigned)(C2 * ((1 << l2) / default_E(l2)));
    void* a2[] = {&B};
    return launch_fn(f2, (unsigned)t2, b2, a2, (size_t)C2 * B.sc * cs, st);
}

int axis_pass(const Workspace& ws, Prec prec, int kind, int logN, const void* in, long long in_e,
              void* out, long long out_e, const Levels& lv, hipStream_t st) {
    if (!ensure_device()) return -1;
    const KTable& kt = ktable(prec);
    const int maxlog = device().maxlog[prec];
    if (kind == K_DFT || kind == K_INVDFT) return c2c_axis(ws, prec, kind, logN, in, in_e, out, out_e, lv, st);
    if (logN == 0) {
        TileArgs A;
        std::memset(&A, 0, sizeof(A));
        fill_levels(A, lv);
        A.in = in;
        A.out = out;
        A.in_e = in_e;
        A.out_e = out_e;
        const unsigned long long ntr = lv.count();
        A.ntrans = (unsigned)ntr;
        const unsigned bs = 256;
        void* args[] = {&A};
        return launch_fn(kt.n1[kind], (unsigned)((ntr + bs - 1) / bs), bs, args, 0, st);
    }
    const int logL = logN - 1;
    if (logL <= maxlog)
        return launch_tile(prec, kind, logL, ws.tw, ws.tw64, in, in_e, out, out_e, lv, -1, 0, st);
    if (!ws.gbuf) {
        set_error("internal: missing two-step scratch");
        return -1;
    }
    if (tune().rlong) {
        const int rc = rlong_axis(ws, prec, kind, logL, in, in_e, out, out_e, lv, st);
        if (rc <= 0) return rc;  // 1: lengths outside the two-pass kernels
    }
    // long real transform: pre-kernel -> two-step complex FFT -> post-kernel
    if (!ws.fbuf) {
        set_error("internal: missing two-step scratch");
        return -1;
    }
    // (with entry tables: the pre-kernel reads through the input table and
    // writes the compact scratch, the post-kernel writes through the output
    // table; the two-step FFT runs on the compact scratch)
    const int L = 1 << logL;
    const unsigned long long ntr = lv.count();
    TileArgs A;
    std::memset(&A, 0, sizeof(A));
    fill_levels(A, lv);
    A.ntrans = (unsigned)ntr;
    A.tw = ws.tw;
    A.in = in;
    A.in_e = in_e;
    A.out = ws.gbuf;
    A.out_off = nullptr;
    int rc = launch_elementwise(kt.pre[kind], A, L, (long long)ntr * L, st);
    if (rc) return rc;
    Levels g;
    if (ntr > 1) g.push((unsigned)ntr, L, L);
    Workspace w2 = ws;
    rc = c2c_axis(w2, prec, kind_inverse_fft(kind) ? K_INVDFT : K_DFT, logL, ws.gbuf, 1, ws.gbuf, 1, g, st);
    if (rc) return rc;
    TileArgs B;
    std::memset(&B, 0, sizeof(B));
    fill_levels(B, lv);
    B.ntrans = (unsigned)ntr;
    B.tw = ws.tw;
    B.tw64 = (const double2*)ws.tw64;  // the post-twiddle is computed in FP64
    B.in = ws.gbuf;
    B.in_off = nullptr;
    B.out = out;
    B.out_e = out_e;
    return launch_elementwise(kt.post[kind], B, L, (long long)ntr * kind_nout(kind, L), st);
}

// ------------------------------------------------------------------ plans
int alloc_twiddles(Workspace& ws, Prec prec, int log_m, hipStream_t st) {
    const long long total = 2LL << log_m;
    HIPCHK(dev_alloc((void**)&ws.tw64, (size_t)total * sizeof(double2)), "hipMalloc twiddles");
    if (prec == PREC_F32) {
        HIPCHK(dev_alloc((void**)&ws.tw, (size_t)total * sizeof(float2)), "hipMalloc twiddles");
    } else {
        ws.tw = ws.tw64;
    }
    ws.tw_log = log_m;
    HIPCHK(launch_twiddles((double2*)ws.tw64, prec == PREC_F32 ? (float2*)ws.tw : nullptr, total, st),
           "twiddle kernel");
    return 0;
}

void free_workspace(Workspace& ws) {
    if (ws.tw && ws.tw != ws.tw64) (void)hipFree(ws.tw);
    if (ws.tw64) (void)hipFree(ws.tw64);
    if (ws.fbuf) (void)hipFree(ws.fbuf);
    if (ws.gbuf) (void)hipFree(ws.gbuf);
    if (ws.cbuf) (void)hipFree(ws.cbuf);
    ws = Workspace();
}

static long long prod_range(const std::vector<int>& v, int a, int b) {
    long long p = 1;
    for (int i = a; i < b; ++i) p *= v[i];
    return p;
}

// kinds used by a family (for sizing)
static void family_kinds(Family f, int* k, int& n) {
    n = 0;
    switch (f) {
        case FAM_DFT: k[n++] = K_DFT; break;
        case FAM_REALDFT: k[n++] = K_REALDFT; break;
        case FAM_T2T3: k[n++] = K_DCT2; break;
        default: k[n++] = K_DCT4; break;
    }
}

void free_plan(Plan* p) {
    if (!p) return;
    if (p->home) (void)hipStreamDestroy(p->home);
    free_workspace(p->ws);
    if (p->tbuf) (void)hipFree(p->tbuf);
    if (p->stage) (void)hipFree(p->stage);
    if (p->hstage) (void)hipHostFree(p->hstage);
    if (p->hflag) (void)hipHostFree(p->hflag);
    for (auto e : p->pev) (void)hipEventDestroy(e);
    p->magic = 0;
    delete p;
}

Plan* make_plan(Family fam, Prec prec, int d, const int* Ns) {
    clear_error();
    if (d < 1 || !Ns) {
        set_error("invalid number of dimensions");
        return nullptr;
    }
    long long P = 1;
    for (int i = 0; i < d; ++i) {
        const int N = Ns[i];
        if (N <= 0 || (N & (N - 1))) {
            set_error("transform lengths must be positive powers of two");
            return nullptr;
        }
        if (P > (1LL << 40) / N) {  // (checked before the product, which could overflow)
            set_error("transform too large");
            return nullptr;
        }
        P *= N;
    }
    if (!ensure_device()) return nullptr;
    Plan* p = new Plan();
    p->fam = fam;
    p->prec = prec;
    p->d = d;
    p->Ns.assign(Ns, Ns + d);
    p->P = P;
    p->dev = device().id;
    auto fail = [&](const char* what) -> Plan* {
        if (what && !*last_error()) set_error(what);
        (void)hipGetLastError();  // the failed call's error stays with this maker
        free_plan(p);
        return nullptr;
    };
    // a blocking stream: its work is ordered after work queued earlier on the
    // legacy default stream, so a synchronous call on device arrays sees what
    // the caller's default-stream kernels wrote into them
    if (hipStreamCreateWithFlags(&p->home, hipStreamDefault) != hipSuccess)
        return fail("hipStreamCreate failed");
    p->stream = p->home;
    const size_t rs = real_size(prec), cs = 2 * rs;
    // twiddle range and scratch over all axis passes of the family
    int kinds[2], nk;
    family_kinds(fam, kinds, nk);
    int twlog = 1;
    size_t fs = 0;
    const long long Nd = Ns[d - 1];
    const long long H = Nd / 2 + 1;
    const long long rows = P / Nd;
    for (int a = 0; a < d; ++a) {
        const int logN = ilog2i(Ns[a]);
        int kind = kinds[0];
        if (fam == FAM_REALDFT && a < d - 1) kind = K_DFT;
        if (const char* why = axis_length_error(prec, kind, logN)) return fail(why);
        twlog = std::max(twlog, needed_tw_log(kind, logN));
        const long long others = (fam == FAM_REALDFT && a < d - 1) ? rows / Ns[a] * H : P / Ns[a];
        const long long extent = (fam == FAM_REALDFT && a < d - 1) ? rows * H : P;
        fs = std::max(fs, axis_scratch_bytes(prec, kind, logN, (unsigned long long)others,
                                             (unsigned long long)extent));
    }
    (void)cs;
    if (alloc_twiddles(p->ws, prec, twlog, p->stream)) return fail(nullptr);
    if (fs) {
        if (dev_alloc((void**)&p->ws.fbuf, fs) != hipSuccess) return fail("hipMalloc scratch failed");
        if (dev_alloc((void**)&p->ws.gbuf, fs) != hipSuccess) return fail("hipMalloc scratch failed");
        p->ws.fs_bytes = fs;
    }
    if (fam == FAM_REALDFT && d > 1) {
        p->tbuf_bytes = (size_t)(rows * H) * 2 * rs;
        if (dev_alloc((void**)&p->tbuf, p->tbuf_bytes) != hipSuccess) return fail("hipMalloc scratch failed");
    }
    if (hipStreamSynchronize(p->stream) != hipSuccess) return fail("twiddle generation failed");
    // launch schedule: timed on the device for large plans, with the
    // family's forward kind standing for all of its kinds
    p->sched = base_tune();
    {
        static const int rep[4] = {K_DFT, K_REALDFT, K_DCT2, K_DCT4};
        TuneRequest r;
        r.prec = prec;
        r.kind = rep[fam];
        r.d = std::min(d, 3);
        for (int k = 0; k < r.d; ++k) r.n[k] = Ns[k];
        r.count = 1;
        r.tbuf = p->tbuf;
        r.solo = true;  // a plan's passes run one after another on its stream
        if (d <= 3) {
            const TuneResult t = autotune_group(p->ws, r, p->stream);
            p->sched = t.sched;
            p->tuned = t.tuned;
        }
        clear_error();
    }
    return p;
}

size_t in_bytes(const Plan& p, int kind) {
    const size_t rs = real_size(p.prec);
    const long long Nd = p.Ns[p.d - 1];
    if (kind == K_DFT || kind == K_INVDFT) return (size_t)p.P * 2 * rs;
    if (kind == K_INVREALDFT) return (size_t)(p.P / Nd * (Nd / 2 + 1)) * 2 * rs;
    return (size_t)p.P * rs;
}

size_t out_bytes(const Plan& p, int kind) {
    const size_t rs = real_size(p.prec);
    const long long Nd = p.Ns[p.d - 1];
    if (kind == K_DFT || kind == K_INVDFT) return (size_t)p.P * 2 * rs;
    if (kind == K_REALDFT) return (size_t)(p.P / Nd * (Nd / 2 + 1)) * 2 * rs;
    return (size_t)p.P * rs;
}

static bool overlap(const void* a, size_t na, const void* b, size_t nb) {
    const char* x = (const char*)a;
    const char* y = (const char*)b;
    return x < y + nb && y < x + na;
}

// levels of the pass along axis a of an array with dims D (item units)
static Levels axis_levels(const std::vector<long long>& Dm, int a, long long& S) {
    const int d = (int)Dm.size();
    S = 1;
    for (int i = a + 1; i < d; ++i) S *= Dm[i];
    long long outer = 1;
    for (int i = 0; i < a; ++i) outer *= Dm[i];
    Levels lv;
    if (outer > 1) lv.push((unsigned)outer, Dm[a] * S, Dm[a] * S);
    if (S > 1) lv.push((unsigned)S, 1, 1);
    return lv;
}

// All passes of `kind` (axes amax down to amin; the first of them reads
// `in`, the others work on `out`) -- see run_entries.
static int run_passes(const Workspace& ws, Prec pr, int kind, int d, const int* Ns, const void* in, void* out,
                      unsigned count, const long long* ioff, const long long* ooff, const long long* foff,
                      void* tbuf, const long long* toff, hipStream_t st, int amax, int amin) {
    std::vector<long long> Dm(Ns, Ns + d);
    long long P = 1;
    for (int i = 0; i < d; ++i) P *= Dm[i];
    const bool tab = ioff != nullptr;
    // batch levels of one pass: the entry level (tables) then the axis levels
    auto mk = [&](const Levels& ax, const long long* it, const long long* ot) {
        Levels lv;
        if (tab) {
            lv.push(count, 0, 0);
            lv.ioff = it;
            lv.ooff = ot;
            lv.foff = foff;
        }
        for (int k = 0; k < ax.n; ++k) lv.push(ax.B[k], ax.is[k], ax.os[k]);
        return lv;
    };
    if (kind != K_REALDFT && kind != K_INVREALDFT) {
        for (int a = amax; a >= amin; --a) {
            long long S;
            Levels ax = axis_levels(Dm, a, S);
            const bool first = (a == amax);
            Levels lv = mk(ax, first ? ioff : ooff, ooff);
            lv.stream_io = (d == 1);
            int rc = axis_pass(ws, pr, kind, ilog2i(Dm[a]), first ? in : (const void*)out, S, out, S, lv, st);
            if (rc) return rc;
        }
        return 0;
    }
    const long long Nd = Dm[d - 1], H = Nd / 2 + 1, rows = P / Nd;
    const int logNd = ilog2i(Nd);
    std::vector<long long> Hm = Dm;
    Hm[d - 1] = H;
    if (kind == K_REALDFT) {
        // real pass along the last axis (into tbuf when tbuf is given: the
        // input and output overlap), then complex passes over the rest
        void* dst1 = tbuf ? tbuf : out;
        const long long* d1off = tbuf ? toff : ooff;
        Levels ax;
        if (rows > 1) ax.push((unsigned)rows, Nd, H);
        int rc = axis_pass(ws, pr, K_REALDFT, logNd, in, 1, dst1, 1, mk(ax, ioff, d1off), st);
        if (rc) return rc;
        for (int a = d - 2; a >= 0; --a) {
            long long S;
            Levels a2 = axis_levels(Hm, a, S);
            const bool first = (a == d - 2);
            const void* src = first ? (const void*)dst1 : (const void*)out;
            Levels lv = mk(a2, first ? d1off : ooff, ooff);
            // the first complex pass sees the real DC and Nyquist columns
            lv.pack = first && H >= 2;
            rc = axis_pass(ws, pr, K_DFT, ilog2i(Dm[a]), src, S, out, S, lv, st);
            if (rc) return rc;
        }
        return 0;
    }
    // inverse real DFT: complex passes over the leading axes into tbuf (the
    // input is preserved), then the real pass along the last axis
    const void* src = in;
    const long long* soff = ioff;
    for (int a = d - 2; a >= 0; --a) {
        long long S;
        Levels a2 = axis_levels(Hm, a, S);
        if (a == 0 && S > H) {
            // h as a level of its own (the packed pair below needs its index)
            a2 = Levels();
            a2.push((unsigned)(S / H), H, H);
            a2.push((unsigned)H, 1, 1);
        }
        const bool first = (a == d - 2);
        Levels lv = mk(a2, first ? ioff : toff, toff);
        // the last complex pass feeds the real pass, which keeps only the
        // real parts of h = 0 and h = H-1
        lv.pack = (a == 0) && H >= 2;
        int rc = axis_pass(ws, pr, K_INVDFT, ilog2i(Dm[a]), first ? in : (const void*)tbuf, S, tbuf, S, lv, st);
        if (rc) return rc;
        src = tbuf;
        soff = toff;
    }
    Levels ax;
    if (rows > 1) ax.push((unsigned)rows, H, Nd);
    return axis_pass(ws, pr, K_INVREALDFT, logNd, src, 1, out, 1, mk(ax, soff, ooff), st);
}

// Multi-dimensional passes kept in the MALL (Tune::mall_mb, MiB; 0 = off):
// the passes of a multi-dimensional transform re-read what the pass before
// them wrote, which the 256 MiB MALL (infinity cache) still holds if the
// data in flight is small enough.  A batch of entries runs all its passes
// on chunks of entries of about mall_mb (in + out) at a time; a single
// transform of d >= 2 dims runs the passes of its axes 1..d-1 on slabs of
// axis 0 of about mall_mb, then the axis-0 pass over the whole array.  The
// later passes then read their input from the MALL, and the earlier output
// they overwrite in place is never written back to HBM.  The real DFT kinds
// (whose passes change the item type) keep one pass per axis.  Measured on
// MI355X it does not pay: 256 x 1024^2 FP32 DCT images 0.977 ms unchunked,
// 0.994-1.14 ms in chunks of 256-32 MiB; 512^3 FP32 1.22 ms against
// 1.31-1.83 ms in slabs of 128-32 MiB -- the column passes are bound by
// their access pattern, not by HBM, and every chunk adds launch ramps and
// tails.  Off by default (a schedule option for other shapes).
int run_entries(const Workspace& ws, Prec pr, int kind, int d, const int* Ns, const void* in, void* out,
                unsigned count, const long long* ioff, const long long* ooff, const long long* foff,
                void* tbuf, const long long* toff, hipStream_t st) {
    const size_t mall = (size_t)std::max(0, tune().mall_mb) << 20;
    if (mall && d >= 2 && kind != K_REALDFT && kind != K_INVREALDFT) {
        long long P = 1;
        for (int i = 0; i < d; ++i) P *= Ns[i];
        const size_t isz = (kind_complex_in(kind) ? 2 : 1) * real_size(pr);
        const size_t osz = (kind_complex_out(kind) ? 2 : 1) * real_size(pr);
        const size_t ebytes = (size_t)P * (isz + osz);
        if (ioff && count > 1) {
            const unsigned per = (unsigned)std::max<size_t>(1, mall / ebytes);
            if (per < count) {
                for (unsigned c0 = 0; c0 < count; c0 += per) {
                    const unsigned cn = std::min(per, count - c0);
                    const int rc = run_passes(ws, pr, kind, d, Ns, in, out, cn, ioff + c0, ooff + c0,
                                              foff ? foff + c0 : nullptr, tbuf, toff ? toff + c0 : nullptr, st,
                                              d - 1, 0);
                    if (rc) return rc;
                }
                return 0;
            }
        } else if (!ioff && count == 1) {
            const size_t plane = ebytes / (size_t)Ns[0];
            const int k = (int)std::max<size_t>(1, mall / plane);
            if (k < Ns[0]) {
                std::vector<int> sub(Ns, Ns + d);
                const long long pitems = P / Ns[0];
                for (int p0 = 0; p0 < Ns[0]; p0 += k) {
                    sub[0] = std::min(k, Ns[0] - p0);
                    const void* ip = static_cast<const char*>(in) + (size_t)p0 * pitems * isz;
                    void* op = static_cast<char*>(out) + (size_t)p0 * pitems * osz;
                    const int rc = run_passes(ws, pr, kind, d, sub.data(), ip, op, 1, nullptr, nullptr, nullptr,
                                              tbuf, nullptr, st, d - 1, 1);
                    if (rc) return rc;
                }
                return run_passes(ws, pr, kind, d, Ns, out, out, 1, nullptr, nullptr, nullptr, tbuf, nullptr, st,
                                  0, 0);
            }
        }
    }
    return run_passes(ws, pr, kind, d, Ns, in, out, count, ioff, ooff, foff, tbuf, toff, st, d - 1, 0);
}

// the full-array temp a call of `kind` uses: the multi-dimensional inverse
// real DFT always, the real DFT when input and output overlap
static bool uses_tbuf(const Plan& p, int kind, const void* in, const void* out) {
    if (p.d < 2) return false;
    if (kind == K_INVREALDFT) return true;
    return kind == K_REALDFT && overlap(in, in_bytes(p, kind), out, out_bytes(p, kind));
}

static size_t round256(size_t b) { return (b + 255) & ~(size_t)255; }

size_t scratch_need(const Plan& p, int kind, const void* in, const void* out) {
    size_t b = p.ws.fs_bytes ? 2 * round256(p.ws.fs_bytes) : 0;
    if (uses_tbuf(p, kind, in, out)) b += round256(p.tbuf_bytes);
    return b;
}

int enqueue_scratch(Plan& p, int kind, const void* in, void* out, hipStream_t st, void* scratch,
                    size_t scratch_bytes) {
    if (family_of(kind) != p.fam) {
        set_error("aux data does not match the transform kind");
        return -1;
    }
    Workspace ws = p.ws;
    void* tbuf = uses_tbuf(p, kind, in, out) ? p.tbuf : nullptr;
    if (scratch) {
        if (scratch_need(p, kind, in, out) > scratch_bytes) {
            set_error("scratch too small for this transform");
            return -1;
        }
        char* b = static_cast<char*>(scratch);
        size_t at = 0;
        if (ws.fs_bytes) {
            ws.fbuf = b;
            ws.gbuf = b + round256(ws.fs_bytes);
            at = 2 * round256(ws.fs_bytes);
        }
        if (tbuf) tbuf = b + at;
    }
    TuneScope scope(&p.sched);
    return run_entries(ws, p.prec, kind, p.d, p.Ns.data(), in, out, 1, nullptr, nullptr, nullptr, tbuf,
                       nullptr, st);
}

int enqueue(Plan& p, int kind, const void* in, void* out, hipStream_t st) {
    return enqueue_scratch(p, kind, in, out, st, nullptr, 0);
}

// largest host array staged through the aux's pinned bounce buffer
constexpr size_t kPinnedStageMax = (size_t)64 << 20;
// pieces of the pipelined bounce copies (at most kMaxPieces of them)
constexpr size_t kPipePiece = (size_t)4 << 20;
constexpr size_t kMaxPieces = kPinnedStageMax / kPipePiece;
static bool host_pipe_enabled() {
    const char* e = std::getenv("MINFFT_HOST_PIPE");
    return !(e && std::atoi(e) == 0);
}
// largest transform (points in all) whose kernels work on the pinned buffer
// directly (one tile pass per axis for every kind at either precision)
constexpr long long kZeroCopyMaxPoints = 4096;

// Wait for the kernels queued on the plan's stream: a one-lane kernel
// queued behind them writes a new value into the plan's completion word and
// the host polls that word, which returns sooner than a stream
// synchronisation (a 1024-point FP64 round trip on device arrays: 23.7
// against 30.0 us).  The poll is bounded: after 2 ms the stream is
// synchronised as well.  Used only when the last operation is a kernel (a
// DMA copy followed by the flag kernel waited longer); MINFFT_HOSTPOLL=0
// turns it off.
static bool host_poll_enabled() {
    const char* e = std::getenv("MINFFT_HOSTPOLL");
    return !(e && std::atoi(e) == 0);
}

static int wait_kernels(Plan& p);

// MINFFT_SIGNAL=0: the completion word always comes from a second launch
static bool host_signal_enabled() {
    const char* e = std::getenv("MINFFT_SIGNAL");
    return !(e && std::atoi(e) == 0);
}

// polls the plan's completion word for `v` (bounded: false after 2 ms)
static bool poll_flag(const Plan& p, unsigned v) {
    const auto t0 = std::chrono::steady_clock::now();
    for (;;) {
        if (__atomic_load_n(p.hflag, __ATOMIC_ACQUIRE) == v) return true;
        if (std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(2)) return false;
#if defined(__x86_64__)
        __builtin_ia32_pause();
#endif
    }
}

// next value for the completion word (0 is its initial value); the word is
// allocated on first use (false: not available)
static bool next_flag_value(Plan& p, unsigned& v) {
    if (!p.hflag) {
        if (hipHostMalloc((void**)&p.hflag, 64, hipHostMallocCoherent) != hipSuccess) {
            (void)hipGetLastError();
            p.hflag = nullptr;
            return false;
        }
        __atomic_store_n(p.hflag, 0u, __ATOMIC_RELAXED);
        p.hseq = 0;
    }
    if (++p.hseq == 0u) ++p.hseq;
    v = p.hseq;
    return true;
}

// true when `kind` on this plan is exactly one tile launch (1D, at most one
// tile long, neither the radix-2-split nor the residue-split kernel), so that
// the launch itself can store the completion word
static bool single_tile_launch(const Plan& p, int kind) {
    if (p.d != 1 || !ensure_device()) return false;
    const int logN = ilog2i(p.Ns[0]);
    const int maxlog = device().maxlog[p.prec];
    if (kind == K_DFT || kind == K_INVDFT) {
        if (logN > maxlog || p.sched.rsplit > 0) return false;
        const int sp2 = p.sched.split2 >= 0 ? p.sched.split2 : (p.prec == PREC_F64 ? 2 : 3);
        if ((sp2 >= 2 && logN == maxlog) || (sp2 >= 3 && logN == maxlog - 1)) return false;
        return true;
    }
    return logN >= 1 && logN - 1 <= maxlog;
}

// Enqueue `kind` and wait for it; a single tile launch stores the
// completion word itself (no second launch), otherwise wait_kernels
static int run_and_wait(Plan& p, int kind, const void* in, void* out) {
    unsigned v = 0;
    if (host_poll_enabled() && host_signal_enabled() && single_tile_launch(p, kind) && next_flag_value(p, v)) {
        HostSignal sig;
        sig.flag = p.hflag;
        sig.value = v;
        t_signal = &sig;
        const int rc = enqueue(p, kind, in, out, p.stream);
        t_signal = nullptr;
        if (rc) {
            (void)hipStreamSynchronize(p.stream);
            return rc;
        }
        if (sig.used) {
            if (poll_flag(p, v)) return 0;
            HIPCHK(hipStreamSynchronize(p.stream), "transform");
            return 0;
        }
        return wait_kernels(p);
    }
    const int rc = enqueue(p, kind, in, out, p.stream);
    if (rc) {
        (void)hipStreamSynchronize(p.stream);
        return rc;
    }
    return wait_kernels(p);
}

static int wait_kernels(Plan& p) {
    unsigned v = 0;
    if (host_poll_enabled() && next_flag_value(p, v)) {
        if (launch_host_flag(p.hflag, v, p.stream) == hipSuccess) {
            if (poll_flag(p, v)) return 0;
        } else {
            (void)hipGetLastError();
        }
    }
    HIPCHK(hipStreamSynchronize(p.stream), "transform");
    return 0;
}

int execute(Plan& p, int kind, const void* in, void* out) {
    clear_error();
    if (family_of(kind) != p.fam) {
        set_error("aux data does not match the transform kind");
        return -1;
    }
    // the aux's device for this call; the caller's current device is
    // restored on return
    struct DeviceScope {
        int prev = -1;
        explicit DeviceScope(int dev) {
            if (hipGetDevice(&prev) != hipSuccess) {
                (void)hipGetLastError();
                prev = -1;
            }
            if (prev != dev) (void)hipSetDevice(dev);
        }
        ~DeviceScope() {
            int now = -1;
            if (prev >= 0 && hipGetDevice(&now) == hipSuccess && now != prev) (void)hipSetDevice(prev);
        }
    } scope(p.dev);
    const size_t ib = in_bytes(p, kind), ob = out_bytes(p, kind);
    const bool in_host = !is_device_pointer(in), out_host = !is_device_pointer(out);
    if (!in_host && !out_host) return run_and_wait(p, kind, in, out);
    const size_t need = std::max(ib, ob);
    if (p.stage_bytes < need) {
        if (p.stage) (void)hipFree(p.stage);
        p.stage = nullptr;
        p.stage_bytes = 0;
        HIPCHK(dev_alloc((void**)&p.stage, need), "hipMalloc staging buffer");
        p.stage_bytes = need;
    }
    // Host arrays up to kPinnedStageMax go through a pinned bounce buffer
    // owned by the aux (a host memcpy plus one DMA each way); larger ones,
    // or when pinned memory is short, are copied from pageable memory
    void* hb = nullptr;
    if (need <= kPinnedStageMax) {
        if (p.hstage_bytes < need) {
            if (p.hstage) (void)hipHostFree(p.hstage);
            p.hstage = nullptr;
            p.hstage_bytes = 0;
            // coherent: the kernels of the zero-copy path below write it
            // without caching, so their results are in host memory when they
            // complete, whichever XCD ran them
            if (hipHostMalloc(&p.hstage, need, hipHostMallocCoherent) == hipSuccess)
                p.hstage_bytes = need;
            else
                (void)hipGetLastError();
        }
        hb = p.hstage;
    }
    if (in_host && out_host && hb && p.P <= kZeroCopyMaxPoints) {
        // short transforms (at most 4096 points in all: one tile pass per
        // axis) run their kernels on the pinned buffer itself, in place (the
        // passes are in-place safe, as for device arrays), with no DMA copy
        std::memcpy(hb, in, ib);
        if (run_and_wait(p, kind, hb, hb)) return -1;
        std::memcpy(out, hb, ob);
        return 0;
    }
    // Through the bounce buffer in pieces: the host copy of one piece
    // overlaps the DMA of the piece before it, in both directions
    // (MINFFT_HOST_PIPE=0: one piece)
    const size_t piece = host_pipe_enabled() ? kPipePiece : (size_t)-1;
    if (in_host && hb) {
        for (size_t off = 0; off < ib; off += piece) {
            const size_t len = std::min(piece, ib - off);
            std::memcpy(static_cast<char*>(hb) + off, static_cast<const char*>(in) + off, len);
            HIPCHK(hipMemcpyAsync(static_cast<char*>(p.stage) + off, static_cast<char*>(hb) + off, len,
                                  hipMemcpyHostToDevice, p.stream),
                   "copy in");
        }
    } else {
        HIPCHK(hipMemcpyAsync(p.stage, in, ib, hipMemcpyDefault, p.stream), "copy in");
    }
    int rc = enqueue(p, kind, p.stage, p.stage, p.stream);
    if (rc) {
        (void)hipStreamSynchronize(p.stream);
        return rc;
    }
    if (out_host && hb) {
        // (ob <= kPinnedStageMax here, so at most kMaxPieces pieces)
        const size_t np = piece >= ob ? 1 : std::min(kMaxPieces, (ob + piece - 1) / piece);
        const size_t pl = np > 1 ? (ob + np - 1) / np : std::max<size_t>(ob, 1);
        while (p.pev.size() < np) {
            hipEvent_t e;
            HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
            p.pev.push_back(e);
        }
        for (size_t k = 0, off = 0; off < ob; ++k, off += pl) {
            const size_t len = std::min(pl, ob - off);
            HIPCHK(hipMemcpyAsync(static_cast<char*>(hb) + off, static_cast<char*>(p.stage) + off, len,
                                  hipMemcpyDeviceToHost, p.stream),
                   "copy out");
            HIPCHK(hipEventRecord(p.pev[k], p.stream), "hipEventRecord");
        }
        for (size_t k = 0, off = 0; off < ob; ++k, off += pl) {
            const size_t len = std::min(pl, ob - off);
            HIPCHK(hipEventSynchronize(p.pev[k]), "transform");
            std::memcpy(static_cast<char*>(out) + off, static_cast<const char*>(hb) + off, len);
        }
        return 0;
    }
    HIPCHK(hipMemcpyAsync(out, p.stage, ob, hipMemcpyDefault, p.stream), "copy out");
    HIPCHK(hipStreamSynchronize(p.stream), "transform");
    return 0;
}

}  // namespace mfx
