// autotune.cpp -- picks the launch schedule of each batch group by timing
// candidates on the device.
//
// A group (entries of one kind and shape) runs as a fixed sequence of axis
// passes; what is left open is the launch schedule (Tune in engine.hpp):
// tile widths of row and column passes, the core mapping of column passes,
// wavefront-local exchanges, XCD tile order, and for long axes the single-pass radix-2-split kernel, the fused
// kernel, its variant and occupancy, and the N = N1*N2 split of the two-step
// decomposition.  The tuner runs the
// group's passes on a bounded subset of its entries (at most ~512 MiB of
// input plus output, placed contiguously in scratch buffers) and walks the
// knobs that apply to the group one at a time (coordinate descent, up to
// three rounds), keeping a value when it is at least 3 % faster.  Results are cached per (precision,
// kind, shape, count bucket) for the process, and in the file named by
// MINFFT_TUNE_CACHE if set (valid for the device and library build named in
// its header line).  MINFFT_AUTOTUNE=0, or a MINFFT_TUNE override,
// turns the tuner off (the process-wide schedule is used).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include <unistd.h>

#include "autotune.hpp"

namespace mfx {

bool autotune_enabled() {
    const char* a = std::getenv("MINFFT_AUTOTUNE");
    if (a && std::atoi(a) == 0) return false;
    return std::getenv("MINFFT_TUNE") == nullptr;
}

namespace {

std::mutex g_mu;
std::map<std::string, TuneResult> g_cache;
bool g_loaded = false;

std::string key_of(Prec prec, int kind, int d, const int* n, unsigned count) {
    int cb = 0;
    while ((1u << (cb + 1)) <= count && cb < 31) ++cb;
    char b[128];
    std::snprintf(b, sizeof(b), "p%d.k%d.d%d.%dx%dx%d.c%d", (int)prec, kind, d, n[0], d > 1 ? n[1] : 1,
                  d > 2 ? n[2] : 1, cb);
    return b;
}

// Cache file: a header line naming the device, its CU count and the
// library build, then one line per group:
//   key tpw cfb cfcore wave xcd e32 e4 split split2 fused wpc falt rsplit coop coopla mall csplit us_default us_best
// A file whose header names another device or build is ignored (and
// replaced on the next save).  Saves rewrite the whole file into a
// temporary one and rename it into place, so a reader never sees a partial
// file; processes saving at the same time keep the last writer's entries.
std::string cache_header() {
    hipDeviceProp_t prop;
    std::string arch = "unknown";
    int ncu = 0;
    if (hipGetDeviceProperties(&prop, device().id) == hipSuccess) {
        arch = prop.gcnArchName;
        ncu = prop.multiProcessorCount;
    }
    for (char& c : arch)
        if (c == ' ') c = '_';
    char b[256];
    std::snprintf(b, sizeof(b), "%s_%s", __DATE__, __TIME__);
    // (__DATE__ pads single-digit days with a space)
    std::string build;
    for (const char* c = b; *c; ++c)
        if (*c != ' ') build += *c;
    std::snprintf(b, sizeof(b), "#minfft-tune v5 arch=%s cus=%d build=%s", arch.c_str(), ncu, build.c_str());
    return b;
}

void load_cache() {
    if (g_loaded) return;
    g_loaded = true;
    const char* f = std::getenv("MINFFT_TUNE_CACHE");
    if (!f) return;
    std::ifstream in(f);
    std::string line;
    if (!std::getline(in, line) || line != cache_header()) return;
    while (std::getline(in, line)) {
        std::istringstream ss(line);
        std::string k;
        TuneResult r;
        r.sched = base_tune();
        Tune& t = r.sched;
        if (ss >> k >> t.tpw >> t.cfb >> t.cfcore >> t.wave >> t.xcd >> t.e32 >> t.e4 >> t.split >> t.split2 >>
            t.fused >> t.fused_wpc >> t.fused_alt >> t.rsplit >> t.coop >> t.coop_la >> t.mall_mb >> t.csplit >>
            r.us_default >> r.us_best) {
            r.tuned = true;
            r.cached = true;
            g_cache[k] = r;
        }
    }
}

// rewrites the cache file with every entry of this process (loaded or tuned)
void save_cache() {
    const char* f = std::getenv("MINFFT_TUNE_CACHE");
    if (!f || !*f) return;
    const std::string tmp = std::string(f) + ".tmp." + std::to_string((long long)getpid());
    {
        std::ofstream out(tmp, std::ios::trunc);
        if (!out) return;
        out << cache_header() << '\n';
        for (const auto& kv : g_cache) {
            const TuneResult& r = kv.second;
            const Tune& t = r.sched;
            out << kv.first << ' ' << t.tpw << ' ' << t.cfb << ' ' << t.cfcore << ' ' << t.wave << ' ' << t.xcd << ' '
                << t.e32 << ' ' << t.e4 << ' ' << t.split << ' ' << t.split2 << ' ' << t.fused << ' ' << t.fused_wpc
                << ' ' << t.fused_alt << ' ' << t.rsplit << ' ' << t.coop << ' ' << t.coop_la << ' ' << t.mall_mb
                << ' ' << t.csplit << ' ' << r.us_default
                << ' ' << r.us_best << '\n';
        }
        if (!out.good()) {
            out.close();
            std::remove(tmp.c_str());
            return;
        }
    }
    if (std::rename(tmp.c_str(), f) != 0) std::remove(tmp.c_str());
}

struct Knob {
    int Tune::*field;
    std::vector<int> values;
};

}  // namespace

// items per entry of `kind` with dims n (complex items for complex sides)
static void entry_items(int kind, int d, const int* n, long long& in_items, long long& out_items) {
    long long P = 1;
    for (int k = 0; k < d; ++k) P *= n[k];
    const long long last = n[d - 1], rows = P / last, H = last / 2 + 1;
    in_items = out_items = P;
    if (kind == K_REALDFT) out_items = rows * H;
    if (kind == K_INVREALDFT) in_items = rows * H;
}

TuneResult autotune_group(const Workspace& ws, const TuneRequest& r, hipStream_t st) {
    TuneResult res;
    res.sched = base_tune();
    if (!autotune_enabled() || r.count == 0) return res;
    const std::string key = key_of(r.prec, r.kind, r.d, r.n, r.count) + (r.solo ? ".solo" : "");
    {
        std::lock_guard<std::mutex> lk(g_mu);
        load_cache();
        auto it = g_cache.find(key);
        if (it != g_cache.end()) return it->second;
    }
    const size_t rs = real_size(r.prec);
    const size_t isz = (kind_complex_in(r.kind) ? 2 : 1) * rs, osz = (kind_complex_out(r.kind) ? 2 : 1) * rs;
    long long ii, oi;
    entry_items(r.kind, r.d, r.n, ii, oi);
    const size_t ebytes = (size_t)ii * isz + (size_t)oi * osz;
    // too little work to measure reliably: keep the defaults
    if (ebytes * r.count < ((size_t)4 << 20)) return res;
    const unsigned cs = (unsigned)std::min<size_t>(r.count, std::max<size_t>(1, ((size_t)512 << 20) / ebytes));

    int maxlog = device().maxlog[r.prec];
    bool long_axis = false, near_limit = false, multi = r.d > 1, short_row = r.n[r.d - 1] <= 1024;
    for (int k = 0; k < r.d; ++k) {
        if (r.n[k] > (1 << maxlog)) long_axis = true;
        if (r.n[k] >= (1 << (maxlog - 1))) near_limit = true;
    }
    const bool fused_ok = r.d == 1 && long_axis && (r.kind == K_DFT || r.kind == K_INVDFT) &&
                          cs >= (unsigned)device().ncu;

    std::vector<Knob> knobs;
    if (fused_ok) {
        knobs.push_back({&Tune::fused, {1, 0}});
        knobs.push_back({&Tune::fused_alt, {0, 1}});
        knobs.push_back({&Tune::fused_wpc, {2, 1, 4}});
        knobs.push_back({&Tune::coop, {-1, 1, 0}});
        knobs.push_back({&Tune::coop_la, {8, 4, 16}});
    }
    if (near_limit) knobs.push_back({&Tune::split2, {-1, 1, 2, 3, 0}});
    if (r.solo && (r.kind == K_DFT || r.kind == K_INVDFT)) {
        // residue-split kernels for the contiguous (last) axis -- only for a
        // plan's only group: timed alone they can win, but beside other
        // groups on concurrent branches their L2-heavy reads and partial-line
        // writes slow the neighbours more than they gain (on the 8192-entry
        // uneven FP32 batch: 490 us per step without them, 529-540 us with
        // the 2^14 and 2^15 groups on them)
        int lg = 0;
        while ((1 << (lg + 1)) <= r.n[r.d - 1]) ++lg;
        std::vector<int> codes = rsplit_codes(r.prec, lg);
        if (!codes.empty()) {
            codes.insert(codes.begin(), 0);
            knobs.insert(knobs.begin(), {&Tune::rsplit, codes});
        }
    }
    if (long_axis) knobs.push_back({&Tune::split, {0, -1, 1}});
    if (multi || long_axis) {
        knobs.push_back({&Tune::cfcore, {-1, 0, 1}});
        knobs.push_back({&Tune::cfb, {128, 64, 256, 32}});
        knobs.push_back({&Tune::xcd, {1, 0}});
    }
    if (multi && (r.kind == K_DFT || r.kind == K_INVDFT)) knobs.push_back({&Tune::csplit, {-1, 0, 1}});
    {
        // radix-32 tiles and radix-2-split halves (FP32): FFT lengths of
        // 2^9 .. 2^(maxlog+1) points (a real kind's FFT is half its length)
        const bool cplx = r.kind == K_DFT || r.kind == K_INVDFT;
        bool mid = false;
        for (int k = 0; k < r.d; ++k) {
            const int fl = (cplx || (r.kind <= K_INVREALDFT && k < r.d - 1)) ? r.n[k] : r.n[k] / 2;
            if (fl >= 512 && fl <= (2 << maxlog)) mid = true;
        }
        if (mid && r.prec == PREC_F32) knobs.push_back({&Tune::e32, {0, 1}});
    }
    knobs.push_back({&Tune::tpw, {256, 128, 512}});
    if (short_row) knobs.push_back({&Tune::wave, {1, 0}});

    void *din = nullptr, *dout = nullptr;
    long long *dio = nullptr, *doo = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    auto cleanup = [&] {
        if (din) (void)hipFree(din);
        if (dout) (void)hipFree(dout);
        if (dio) (void)hipFree(dio);
        if (doo) (void)hipFree(doo);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
    };
    std::vector<long long> hio(cs), hoo(cs);
    for (unsigned i = 0; i < cs; ++i) {
        hio[i] = (long long)i * ii;
        hoo[i] = (long long)i * oi;
    }
    if (dev_alloc((void**)&din, (size_t)cs * ii * isz) != hipSuccess ||
        dev_alloc((void**)&dout, (size_t)cs * oi * osz) != hipSuccess ||
        dev_alloc((void**)&dio, cs * sizeof(long long)) != hipSuccess ||
        dev_alloc((void**)&doo, cs * sizeof(long long)) != hipSuccess ||
        hipMemcpy(dio, hio.data(), cs * sizeof(long long), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(doo, hoo.data(), cs * sizeof(long long), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemsetAsync(din, 0, (size_t)cs * ii * isz, st) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
        hipEventCreate(&e1) != hipSuccess) {
        (void)hipGetLastError();
        cleanup();
        return res;  // no room to tune: defaults
    }

    auto measure = [&](const Tune& t) -> float {
        TuneScope scope(&t);
        auto once = [&]() {
            if (cs == 1)  // a single transform, as the batch runs a one-entry group
                return run_entries(ws, r.prec, r.kind, r.d, r.n, din, dout, 1, nullptr, nullptr, nullptr, r.tbuf,
                                   nullptr, st);
            return run_entries(ws, r.prec, r.kind, r.d, r.n, din, dout, cs, dio, doo, r.d_foff, r.tbuf, r.d_toff, st);
        };
        if (once()) {
            clear_error();
            return INFINITY;
        }
        constexpr int reps = 3;
        if (hipEventRecord(e0, st) != hipSuccess) return INFINITY;
        for (int i = 0; i < reps; ++i)
            if (once()) {
                clear_error();
                (void)hipStreamSynchronize(st);
                return INFINITY;
            }
        if (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) return INFINITY;
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) return INFINITY;
        return ms * 1000.f / reps;
    };

    Tune best = base_tune();
    // The GPU leaves its idle power state only after ~10 ms of load, and
    // candidates timed during that ramp would look faster than the default
    // timed first: run the default schedule for ~20 ms before timing.
    {
        TuneScope scope(&best);
        const auto t0 = std::chrono::steady_clock::now();
        for (int i = 0; i < 1000; ++i) {
            if (run_entries(ws, r.prec, r.kind, r.d, r.n, din, dout, cs, dio, doo, r.d_foff, r.tbuf, r.d_toff,
                            st)) {
                clear_error();
                break;
            }
            if ((i & 3) == 3) {
                (void)hipStreamSynchronize(st);
                if (std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(20)) break;
            }
        }
        (void)hipStreamSynchronize(st);
    }
    float tbest = measure(best);
    float tdef = tbest;
    if (std::isfinite(tbest)) {
        // coordinate descent, repeated while a round changes something
        // (knobs interact: the radix-2-split level and the radix-32 choice)
        for (int round = 0; round < 3; ++round) {
            bool changed = false;
            for (const Knob& k : knobs) {
                for (int v : k.values) {
                    if (best.*(k.field) == v) continue;
                    Tune c = best;
                    c.*(k.field) = v;
                    const float t = measure(c);
                    if (t < 0.97f * tbest) {
                        best = c;
                        tbest = t;
                        changed = true;
                    }
                }
            }
            if (!changed) break;
        }
        // the clocks may drift during the descent: time the default and the
        // choice once more, back to back, and keep the choice only if it
        // still wins by 3 %
        const Tune def = base_tune();
        if (std::memcmp(&def, &best, sizeof(Tune)) != 0) {
            const float td = measure(def), tb = measure(best);
            if (std::isfinite(td) && std::isfinite(tb)) {
                tdef = td;
                tbest = tb;
                if (!(tb < 0.97f * td)) {
                    best = def;
                    tbest = td;
                }
            }
        }
        res.sched = best;
        res.tuned = true;
        res.us_default = tdef;
        res.us_best = tbest;
    }
    (void)hipStreamSynchronize(st);
    cleanup();
    if (res.tuned) {
        std::lock_guard<std::mutex> lk(g_mu);
        g_cache[key] = res;
        save_cache();
    }
    return res;
}

std::string describe_tune(const Tune& t) {
    char b[256];
    std::snprintf(b, sizeof(b),
                  "tpw=%d cfb=%d cfcore=%d wave=%d xcd=%d e32=%d split=%d split2=%d fused=%d wpc=%d falt=%d rsplit=%d "
                  "coop=%d coopla=%d mall=%d csplit=%d",
                  t.tpw, t.cfb, t.cfcore, t.wave, t.xcd, t.e32, t.split, t.split2, t.fused, t.fused_wpc, t.fused_alt,
                  t.rsplit, t.coop, t.coop_la, t.mall_mb, t.csplit);
    return b;
}

}  // namespace mfx
