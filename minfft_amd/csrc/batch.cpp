// batch.cpp -- minfft_ext.h: error channel, asynchronous enqueue and the
// batch extension.  A batch is a list of independent entries (kind, dims,
// byte offsets); its semantics are those of calling the single-transform
// routine on every entry (minfft.c:287-290 for the complex DFT, 735-738 /
// 931-934 for DCT-2 / DCT-4, and so on).
//
// Execution: entries with identical (kind, dims) form a group; each group is
// executed as the same per-axis passes as a single transform, with one extra
// batch level whose base addresses come from device tables of per-entry
// offsets, so a group of any size costs one launch per axis (up to three
// for an axis longer than one tile: the two steps of the N = N1*N2
// decomposition plus a real kind's pre or post pass, unless the fused or the
// radix-2-split kernel covers it in one).  The short 1D complex groups
// (2^4 .. 2^12 points) of one direction instead share one persistent launch
// (multi.hpp).  Each group's launch schedule is picked when the plan is made
// by timing candidates on the device (autotune.cpp).  The launches of a plan
// with several units of work are captured once into a HIP graph per (in,
// out, stream), their units on concurrent branches, and replayed; a plan
// with one unit launches straight into the stream (minfft_batch_run).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

#include "autotune.hpp"
#include "engine.hpp"
#include "minfft_ext.h"
#include "multi.hpp"

using namespace mfx;

struct BatchGroup {
    int kind = 0;
    int d = 1;
    int n[3] = {1, 1, 1};
    std::vector<long long> ioff, ooff, foff, toff;  // host copies (items)
    long long* d_ioff = nullptr;
    long long* d_ooff = nullptr;
    long long* d_foff = nullptr;
    long long* d_toff = nullptr;
    size_t tbytes = 0;     // per-group temp (inverse real DFT, d > 1)
    size_t fbytes = 0;     // per-group two-step scratch
    void* tbuf = nullptr;
    TuneResult tune;       // launch schedule of the group's passes
    bool in_multi = false; // runs inside a persistent short-length launch
};

// One persistent launch for the short 1D complex groups of one direction
// (multi.hpp); `args` gets the buffers at enqueue.
struct MultiLaunch {
    bool inv = false;
    std::vector<int> members;  // group indices
    MultiArgs args;
    unsigned grid = 0;
    long long bytes = 0;
};

struct minfft_batch_plan {
    Prec prec = PREC_F32;
    int prec_req = 1;
    std::vector<BatchGroup> groups;
    Workspace ws;
    hipStream_t stream = nullptr;
    // graph cache
    const void* g_in = nullptr;
    void* g_out = nullptr;
    hipStream_t g_st = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    // persistent launches of the short 1D complex groups, and their item
    // counters (one per launch)
    std::vector<MultiLaunch> multi;
    unsigned* d_counter = nullptr;
    // device staging of minfft_batch's host arrays (grown on demand)
    void* st_in = nullptr;
    void* st_out = nullptr;
    size_t st_in_bytes = 0, st_out_bytes = 0;
    // branch streams/events for concurrent groups inside the captured graph
    std::vector<hipStream_t> side;
    std::vector<hipEvent_t> ev;
    // bytes of `in` / `out` the entries address (largest offset + extent)
    size_t in_extent = 0, out_extent = 0;
    // the last (in, out) pair found to be device memory
    const void* ok_in = nullptr;
    const void* ok_out = nullptr;
};

// short 1D complex groups share one persistent launch per direction
// (MINFFT_MULTI=0: one launch per group)
static bool multi_enabled() {
    const char* e = std::getenv("MINFFT_MULTI");
    return !(e && std::atoi(e) == 0);
}

// Form of the short-length launch.  Alone (the plan's only unit of work)
// one work-group per item is faster -- no counter, no reset launch, the
// hardware spreads the work-groups (5220 entries of 2^6..2^12, FP32:
// 21.8-22.1 against 36.5-36.6 us); beside other groups the persistent
// form is, because its ~1000 resident work-groups leave CUs to the long
// groups on the critical path (the default 8192-entry step: 409.4 against
// 413.5-415.6 us).  MINFFT_MULTI=1 / 2 force the persistent / per-item form.
static bool multi_per_item(size_t units) {
    const char* e = std::getenv("MINFFT_MULTI");
    const int v = e ? std::atoi(e) : -1;
    if (v == 1) return false;
    if (v == 2) return true;
    return units <= 1;
}

// groups run on up to this many concurrent graph branches (default 8: in
// practice one branch per unit of work; the default batch's five units ran
// 0.409 ms per step on five branches against 0.417 ms on four)
static int branch_count() {
    const char* e = std::getenv("MINFFT_BRANCHES");
    int v = e ? std::atoi(e) : 8;
    return v < 1 ? 1 : (v > 8 ? 8 : v);
}

// MINFFT_GRAPH: 1 = always replay a captured graph, 0 = always launch
// directly into the stream, unset = by plan size
static int graph_mode() {
    const char* e = std::getenv("MINFFT_GRAPH");
    return e ? (std::atoi(e) != 0 ? 1 : 0) : -1;
}

static size_t item_in_bytes(int kind, Prec p) { return (kind_complex_in(kind) ? 2 : 1) * real_size(p); }
static size_t item_out_bytes(int kind, Prec p) { return (kind_complex_out(kind) ? 2 : 1) * real_size(p); }

// size limits of a batch: an entry as large as the plain makers accept, and
// bounds on offsets and the points of all entries that no device buffer
// reaches, so that the extent arithmetic below cannot overflow
constexpr long long kMaxEntryPoints = 1LL << 40;
constexpr long long kMaxOffset = 1LL << 50;
constexpr long long kMaxBatchPoints = 1LL << 44;

static int lg2(long long x) {
    int l = 0;
    while ((1LL << l) < x) ++l;
    return l;
}

static long long group_bytes(const BatchGroup& g, Prec pr) {
    long long P = 1;
    for (int k = 0; k < g.d; ++k) P *= g.n[k];
    return P * (long long)g.ioff.size() * (long long)real_size(pr);
}

extern "C" {

const char* minfft_last_error(void) { return mfx::last_error(); }

int minfft_enqueue(int kind, const void* in, void* out, const void* aux, void* stream) {
    clear_error();
    Plan* p = reinterpret_cast<Plan*>(const_cast<void*>(aux));
    if (!p || p->magic != 0x6d667870u) {
        set_error("invalid aux data");
        return -1;
    }
    if (kind < 0 || kind >= K_COUNT) {
        set_error("invalid kind");
        return -1;
    }
    // the kernels dereference in/out directly: host memory would fault the
    // GPU (no XNACK), so it is refused here
    if (!is_device_pointer(in) || !is_device_pointer(out)) {
        set_error("minfft_enqueue needs device (or managed) pointers for in and out");
        return -1;
    }
    hipStream_t st = stream ? reinterpret_cast<hipStream_t>(stream) : p->stream;
    return enqueue(*p, kind, in, out, st);
}

int minfft_set_stream(const void* aux, void* stream) {
    clear_error();
    Plan* p = reinterpret_cast<Plan*>(const_cast<void*>(aux));
    if (!p || p->magic != 0x6d667870u) {
        set_error("invalid aux data");
        return -1;
    }
    if (!stream) {
        p->stream = p->home;
        return 0;
    }
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    // a stream of another GPU would run the aux's kernels on memory that GPU
    // does not hold
    hipDevice_t sd = -1;
    if (hipStreamGetDevice(st, &sd) != hipSuccess || sd != p->dev) {
        (void)hipGetLastError();
        set_error("stream does not belong to the aux's device");
        return -1;
    }
    p->stream = st;
    return 0;
}

void minfft_free_batch(minfft_batch_plan* p) {
    if (!p) return;
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    for (auto& g : p->groups) {
        if (g.d_ioff) (void)hipFree(g.d_ioff);
        if (g.d_ooff) (void)hipFree(g.d_ooff);
        if (g.d_foff) (void)hipFree(g.d_foff);
        if (g.d_toff) (void)hipFree(g.d_toff);
        if (g.tbuf) (void)hipFree(g.tbuf);
    }
    if (p->d_counter) (void)hipFree(p->d_counter);
    if (p->st_in) (void)hipFree(p->st_in);
    if (p->st_out) (void)hipFree(p->st_out);
    free_workspace(p->ws);
    for (auto s : p->side) (void)hipStreamDestroy(s);
    for (auto e : p->ev) (void)hipEventDestroy(e);
    if (p->stream) (void)hipStreamDestroy(p->stream);
    delete p;
}

minfft_batch_plan* minfft_mkaux_batch(int count, const minfft_batch_entry* e, int prec) {
    clear_error();
    if (count <= 0 || !e) {
        set_error("empty batch");
        return nullptr;
    }
    if (prec != MINFFT_FP64 && prec != MINFFT_FP32 && prec != MINFFT_SPLIT_FP16) {
        set_error("unknown precision");
        return nullptr;
    }
    minfft_batch_plan* p = new minfft_batch_plan();
    p->prec = prec == MINFFT_FP64 ? PREC_F64 : PREC_F32;  // split-FP16 keeps FP32 data
    p->prec_req = prec;
    auto fail = [&](const char* msg) -> minfft_batch_plan* {
        if (msg) set_error(msg);
        (void)hipGetLastError();  // the failed call's error stays with this plan
        minfft_free_batch(p);
        return nullptr;
    };
    std::map<std::tuple<int, int, int, int, int>, int> idx;
    int twlog = 1;
    size_t fs_max = 0;
    long long total_points = 0;
    for (int i = 0; i < count; ++i) {
        const minfft_batch_entry& x = e[i];
        if (x.kind < 0 || x.kind >= K_COUNT) return fail("invalid entry kind");
        if (x.d < 1 || x.d > 3) return fail("entry dimensions must be 1..3");
        int n[3] = {1, 1, 1};
        long long P = 1;
        for (int k = 0; k < x.d; ++k) {
            if (x.n[k] <= 0 || (x.n[k] & (x.n[k] - 1))) return fail("entry lengths must be powers of two");
            n[k] = x.n[k];
            if (P > kMaxEntryPoints / n[k]) return fail("entry too large");  // (as the plain makers)
            P *= n[k];
        }
        const size_t ib = item_in_bytes(x.kind, p->prec), ob = item_out_bytes(x.kind, p->prec);
        if (x.in_offset < 0 || x.out_offset < 0 || x.in_offset % (long long)ib || x.out_offset % (long long)ob)
            return fail("entry offsets must be non-negative multiples of the item size");
        if (x.in_offset > kMaxOffset || x.out_offset > kMaxOffset)
            return fail("entry offset beyond any device buffer");
        total_points += P;
        if (total_points > kMaxBatchPoints) return fail("batch too large");
        auto key = std::make_tuple(x.kind, x.d, n[0], n[1], n[2]);
        auto it = idx.find(key);
        int gi;
        if (it == idx.end()) {
            gi = (int)p->groups.size();
            idx[key] = gi;
            BatchGroup g;
            g.kind = x.kind;
            g.d = x.d;
            std::memcpy(g.n, n, sizeof(n));
            p->groups.push_back(g);
        } else {
            gi = it->second;
        }
        BatchGroup& g = p->groups[gi];
        g.ioff.push_back(x.in_offset / (long long)ib);
        g.ooff.push_back(x.out_offset / (long long)ob);
        long long ii = P, oi = P;
        if (x.kind == K_REALDFT) oi = P / n[x.d - 1] * (n[x.d - 1] / 2 + 1);
        if (x.kind == K_INVREALDFT) ii = P / n[x.d - 1] * (n[x.d - 1] / 2 + 1);
        p->in_extent = std::max(p->in_extent, (size_t)x.in_offset + (size_t)ii * ib);
        p->out_extent = std::max(p->out_extent, (size_t)x.out_offset + (size_t)oi * ob);
    }
    if (!ensure_device()) return fail(nullptr);  // (entries are checked first, without a device)
    p->ws = Workspace();
    // blocking: ordered after the legacy default stream's earlier work (as the plain aux's stream)
    if (hipStreamCreateWithFlags(&p->stream, hipStreamDefault) != hipSuccess) return fail("hipStreamCreate failed");
    const size_t cs = 2 * real_size(p->prec);
    for (auto& g : p->groups) {
        const unsigned cnt = (unsigned)g.ioff.size();
        long long P = 1;
        for (int k = 0; k < g.d; ++k) P *= g.n[k];
        const long long Nd = g.n[g.d - 1], H = Nd / 2 + 1, rows = P / Nd;
        const bool realfam = g.kind == K_REALDFT || g.kind == K_INVREALDFT;
        // per-entry complex extent of the arrays the complex passes touch
        const long long cext = realfam ? rows * H : P;
        size_t fs = 0;
        for (int a = 0; a < g.d; ++a) {
            int kind = g.kind;
            if (realfam && a < g.d - 1) kind = K_DFT;
            if (const char* why = axis_length_error(p->prec, kind, lg2(g.n[a]))) return fail(why);
            twlog = std::max(twlog, needed_tw_log(kind, lg2(g.n[a])));
            const long long others = (realfam && a < g.d - 1) ? rows / g.n[a] * H : P / g.n[a];
            fs = std::max(fs, axis_scratch_bytes(p->prec, kind, lg2(g.n[a]), (unsigned long long)others * cnt,
                                                 (unsigned long long)cext * cnt));
        }
        g.fbytes = fs;
        fs_max = std::max(fs_max, fs);
        g.foff.resize(cnt);
        for (unsigned i = 0; i < cnt; ++i) g.foff[i] = (long long)i * cext;
        if (g.kind == K_INVREALDFT && g.d > 1) {
            g.tbytes = (size_t)cnt * (size_t)(rows * H) * cs;
            g.toff.resize(cnt);
            for (unsigned i = 0; i < cnt; ++i) g.toff[i] = (long long)i * rows * H;
            if (dev_alloc((void**)&g.tbuf, g.tbytes) != hipSuccess) return fail("hipMalloc batch temp failed");
        }
        // the group's offset tables on the device (nullptr: done, else the reason)
        auto up = [&](std::vector<long long>& h, long long*& dptr) -> const char* {
            if (h.empty()) return nullptr;
            if (dev_alloc((void**)&dptr, h.size() * sizeof(long long)) != hipSuccess)
                return "hipMalloc batch tables failed";
            if (hipMemcpy(dptr, h.data(), h.size() * sizeof(long long), hipMemcpyHostToDevice) != hipSuccess)
                return "uploading batch tables failed";
            return nullptr;
        };
        for (auto t : {std::make_pair(&g.ioff, &g.d_ioff), std::make_pair(&g.ooff, &g.d_ooff),
                       std::make_pair(&g.foff, &g.d_foff), std::make_pair(&g.toff, &g.d_toff)})
            if (const char* why = up(*t.first, *t.second)) return fail(why);
    }
    // largest groups first
    std::stable_sort(p->groups.begin(), p->groups.end(), [](const BatchGroup& a, const BatchGroup& b) {
        long long pa = 1, pb = 1;
        for (int k = 0; k < a.d; ++k) pa *= a.n[k];
        for (int k = 0; k < b.d; ++k) pb *= b.n[k];
        return pa * (long long)a.ioff.size() > pb * (long long)b.ioff.size();
    });
    if (alloc_twiddles(p->ws, p->prec, twlog, p->stream)) return fail(nullptr);
    p->ws.split16 = (prec == MINFFT_SPLIT_FP16) ? 1 : 0;
    if (fs_max) {
        if (dev_alloc((void**)&p->ws.fbuf, fs_max) != hipSuccess) return fail("hipMalloc batch scratch failed");
        if (dev_alloc((void**)&p->ws.gbuf, fs_max) != hipSuccess) return fail("hipMalloc batch scratch failed");
        p->ws.fs_bytes = fs_max;
        // counters of the XCD-cooperative kernel: 8 queue heads and one word
        // per entry of the largest group
        size_t maxcnt = 0;
        for (const auto& g : p->groups) maxcnt = std::max(maxcnt, g.ioff.size());
        const size_t cb = (8 + maxcnt) * sizeof(unsigned);
        if (dev_alloc((void**)&p->ws.cbuf, cb) != hipSuccess) return fail("hipMalloc batch counters failed");
        p->ws.cbuf_bytes = cb;
    }
    if (hipStreamSynchronize(p->stream) != hipSuccess) return fail("twiddle generation failed");
    // short 1D complex groups (lengths 2^4 .. 2^12): one persistent launch
    // per direction, items numbered largest length first
    const int mgrid = multi_enabled() ? multi_grid(p->prec) : 0;
    for (int inv = 0; inv < 2 && mgrid > 0; ++inv) {
        std::vector<int> mem;
        for (int gi = 0; gi < (int)p->groups.size(); ++gi) {
            const BatchGroup& g = p->groups[gi];
            const int lg = lg2(g.n[0]);
            // (split-FP16 plans keep their 512-point groups on the matrix-core tile)
            if (g.d == 1 && g.kind == (inv ? K_INVDFT : K_DFT) && lg >= 4 && lg <= 12 &&
                !(p->ws.split16 && lg == 9) && (int)mem.size() < kMultiMaxGroups)
                mem.push_back(gi);
        }
        if (mem.empty()) continue;
        std::stable_sort(mem.begin(), mem.end(),
                         [&](int a, int b) { return p->groups[a].n[0] > p->groups[b].n[0]; });
        MultiLaunch ml;
        ml.inv = inv != 0;
        ml.members = mem;
        std::memset(&ml.args, 0, sizeof(ml.args));
        unsigned long long items = 0;
        for (int k = 0; k < (int)mem.size(); ++k) {
            BatchGroup& g = p->groups[mem[k]];
            g.in_multi = true;
            const unsigned cnt = (unsigned)g.ioff.size();
            const unsigned per = (unsigned)(kMultiThreads * 16 / g.n[0]);  // transforms per item
            MultiGroup& mg = ml.args.g[k];
            mg.in_off = g.d_ioff;
            mg.out_off = g.d_ooff;
            mg.count = cnt;
            mg.item0 = (unsigned)items;
            mg.logL = lg2(g.n[0]);
            items += (cnt + per - 1) / per;
            ml.bytes += group_bytes(g, p->prec);
        }
        if (items > 0x7fffffffull) return fail("batch too large");
        ml.args.ngroups = (int)mem.size();
        ml.args.total = (unsigned)items;
        ml.args.tw = p->ws.tw;
        ml.grid = (unsigned)std::min<unsigned long long>(items, (unsigned long long)mgrid);
        p->multi.push_back(ml);
    }
    if (!p->multi.empty()) {
        if (dev_alloc((void**)&p->d_counter, 2 * sizeof(unsigned)) != hipSuccess)
            return fail("hipMalloc batch counter failed");
        for (size_t k = 0; k < p->multi.size(); ++k) p->multi[k].args.counter = p->d_counter + k;
    }
    for (auto& g : p->groups) {
        if (g.in_multi) continue;  // not launched on its own: nothing to tune
        TuneRequest r;
        r.prec = p->prec;
        r.kind = g.kind;
        r.d = g.d;
        std::memcpy(r.n, g.n, sizeof(r.n));
        r.count = (unsigned)g.ioff.size();
        r.d_foff = g.d_foff;
        r.tbuf = g.tbuf;
        r.d_toff = g.d_toff;
        r.solo = p->groups.size() == 1;
        g.tune = autotune_group(p->ws, r, p->stream);
    }
    clear_error();
    return p;
}


// Enqueue every group.  With branches > 1 the groups are spread (largest
// first, onto the least-loaded branch) over side streams forked from and
// joined back into `st`, so that in a captured graph they become parallel
// branches: short groups fill the tails of the long ones.
static int enqueue_batch(minfft_batch_plan* p, const void* in, void* out, hipStream_t st, int branches) {
    // units of work, largest first: the groups launched on their own and
    // the persistent short-length launches (index -1 - k)
    std::vector<std::pair<long long, int>> units;
    for (int gi = 0; gi < (int)p->groups.size(); ++gi)
        if (!p->groups[gi].in_multi) units.push_back({group_bytes(p->groups[gi], p->prec), gi});
    // (a persistent launch counts three times its bytes: on the default batch
    // it then gets a branch of its own and starts early instead of trailing
    // the 2^13 group, 0.424 against 0.428 ms per step over six A/B runs)
    for (int k = 0; k < (int)p->multi.size(); ++k) units.push_back({p->multi[k].bytes * 3, -1 - k});
    std::stable_sort(units.begin(), units.end(), [](const std::pair<long long, int>& a,
                                                    const std::pair<long long, int>& b) { return a.first > b.first; });
    auto run_unit = [&](int u, hipStream_t s) -> int {
        if (u < 0) {
            MultiLaunch& ml = p->multi[-1 - u];
            ml.args.in = in;
            ml.args.out = out;
            return launch_multi(p->prec, ml.inv, ml.args, ml.grid, multi_per_item(units.size()), s);
        }
        BatchGroup& g = p->groups[u];
        TuneScope scope(&g.tune.sched);
        if (g.ioff.size() == 1) {
            // one entry: a single transform at known addresses (no tables)
            const void* ip = static_cast<const char*>(in) + g.ioff[0] * (long long)item_in_bytes(g.kind, p->prec);
            void* op = static_cast<char*>(out) + g.ooff[0] * (long long)item_out_bytes(g.kind, p->prec);
            return run_entries(p->ws, p->prec, g.kind, g.d, g.n, ip, op, 1, nullptr, nullptr, nullptr, g.tbuf,
                               nullptr, s);
        }
        return run_entries(p->ws, p->prec, g.kind, g.d, g.n, in, out, (unsigned)g.ioff.size(), g.d_ioff, g.d_ooff,
                           g.d_foff, g.tbuf, g.d_toff, s);
    };
    const int nb = std::min<int>(branches, (int)units.size());
    if (nb <= 1) {
        for (const auto& u : units)
            if (int rc = run_unit(u.second, st)) return rc;
        return 0;
    }
    while ((int)p->side.size() < nb) {
        hipStream_t s;
        if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) {
            set_error("hipStreamCreate failed");
            return -1;
        }
        p->side.push_back(s);
    }
    while ((int)p->ev.size() < nb + 1) {
        hipEvent_t e;
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
            set_error("hipEventCreate failed");
            return -1;
        }
        p->ev.push_back(e);
    }
    // fork
    if (hipEventRecord(p->ev[nb], st) != hipSuccess) return -1;
    for (int b = 0; b < nb; ++b)
        if (hipStreamWaitEvent(p->side[b], p->ev[nb], 0) != hipSuccess) return -1;
    // units are captured largest first; on the default batch that puts the
    // 2^16 group's launch first, and every other order measured was 20-33 us
    // slower per step (the first-captured branch starts 15-20 us before the
    // others)
    std::vector<long long> load(nb, 0);
    for (const auto& u : units) {
        int b = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        if (u.second >= 0) {
            // the two-step scratch is shared: groups that use it stay on
            // branch 0 (1D complex groups whose schedule is single-pass do not)
            const BatchGroup& g = p->groups[u.second];
            const bool single = g.d == 1 && (g.kind == K_DFT || g.kind == K_INVDFT) &&
                                c2c_single_pass(p->prec, lg2(g.n[0]), g.tune.sched, in == out);
            if (g.fbytes && !single) b = 0;
        }
        load[b] += u.first;
        if (int rc = run_unit(u.second, p->side[b])) return rc;
    }
    // join
    for (int b = 0; b < nb; ++b) {
        if (hipEventRecord(p->ev[b], p->side[b]) != hipSuccess) return -1;
        if (hipStreamWaitEvent(st, p->ev[b], 0) != hipSuccess) return -1;
    }
    return 0;
}

int minfft_batch_run(minfft_batch_plan* p, const void* in, void* out, void* stream) {
    clear_error();
    if (!p) {
        set_error("invalid batch plan");
        return -1;
    }
    // the kernels dereference in/out directly (host memory would fault the
    // GPU): checked once per (in, out) pair, i.e. when a graph is captured
    if (!(in == p->ok_in && out == p->ok_out)) {
        if (!is_device_pointer(in) || !is_device_pointer(out)) {
            set_error("minfft_batch_run needs device (or managed) pointers for in and out");
            return -1;
        }
        p->ok_in = in;
        p->ok_out = out;
    }
    hipStream_t st = stream ? reinterpret_cast<hipStream_t>(stream) : p->stream;
    // a plan with one unit of work (one group, or only the persistent
    // launch) has no branches to run concurrently: its launches go straight
    // into the stream, which costs less than replaying a graph (a small
    // group: 6-7 against 16-18 us per call); plans with several units replay
    // the captured graph, whose branches run concurrently at a lower cost
    // than eager cross-stream synchronisation
    int units = (int)p->multi.size();
    for (const BatchGroup& g : p->groups) units += g.in_multi ? 0 : 1;
    const int gm = graph_mode();
    if (gm == 0 || (gm < 0 && units <= 1)) return enqueue_batch(p, in, out, st, branch_count());
    // a caller capturing `st` into a graph of its own gets the launches in
    // that capture (the branches become branches of the caller's graph):
    // a capture of our own cannot begin on a stream that is capturing
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(st, &cap) != hipSuccess) {
        (void)hipGetLastError();
        set_error("hipStreamIsCapturing failed");
        return -1;
    }
    if (cap != hipStreamCaptureStatusNone) return enqueue_batch(p, in, out, st, branch_count());
    if (!(p->exec && p->g_in == in && p->g_out == out && p->g_st == st)) {
        if (p->exec) (void)hipGraphExecDestroy(p->exec);
        if (p->graph) (void)hipGraphDestroy(p->graph);
        p->exec = nullptr;
        p->graph = nullptr;
        if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess) {
            set_error("hipStreamBeginCapture failed");
            return -1;
        }
        int rc = enqueue_batch(p, in, out, st, branch_count());
        hipGraph_t g = nullptr;
        hipError_t ec = hipStreamEndCapture(st, &g);
        if (rc || ec != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            if (!rc) set_error("hipStreamEndCapture failed");
            return -1;
        }
        if (hipGraphInstantiate(&p->exec, g, nullptr, nullptr, 0) != hipSuccess) {
            (void)hipGraphDestroy(g);
            set_error("hipGraphInstantiate failed");
            return -1;
        }
        p->graph = g;
        p->g_in = in;
        p->g_out = out;
        p->g_st = st;
    }
    if (hipGraphLaunch(p->exec, st) != hipSuccess) {
        set_error("hipGraphLaunch failed");
        return -1;
    }
    return 0;
}

int minfft_batch_schedule(minfft_batch_plan* p, int group, const char* spec) {
    clear_error();
    if (!p || group < 0 || group >= (int)p->groups.size()) {
        set_error("invalid batch plan or group");
        return -1;
    }
    TuneResult r;
    r.sched = base_tune();
    if (spec && *spec && !parse_tune(spec, r.sched)) {
        set_error("unknown schedule key");
        return -1;
    }
    r.tuned = false;
    p->groups[group].tune = r;
    // launches are recorded at capture: drop the cached graph
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    p->exec = nullptr;
    p->graph = nullptr;
    p->g_in = nullptr;
    p->g_out = nullptr;
    p->g_st = nullptr;
    return 0;
}

int minfft_batch_describe(const minfft_batch_plan* p, char* buf, int len) {
    clear_error();
    if (!p) {
        set_error("invalid batch plan");
        return -1;
    }
    static const char* kname[K_COUNT] = {"dft", "invdft", "realdft", "invrealdft", "dct2",
                                         "dst2", "dct3", "dst3", "dct4", "dst4"};
    std::string s;
    char line[320];
    for (size_t i = 0; i < p->groups.size(); ++i) {
        const BatchGroup& g = p->groups[i];
        std::string dims = std::to_string(g.n[0]);
        for (int k = 1; k < g.d; ++k) dims += "x" + std::to_string(g.n[k]);
        if (g.in_multi)
            std::snprintf(line, sizeof(line),
                          "group %zu: %s %s count=%zu persistent: one launch with the other short %s groups\n", i,
                          kname[g.kind], dims.c_str(), g.ioff.size(), kname[g.kind]);
        else
            std::snprintf(line, sizeof(line), "group %zu: %s %s count=%zu %s: %s default_us=%.1f tuned_us=%.1f\n",
                          i, kname[g.kind], dims.c_str(), g.ioff.size(), g.tune.tuned ? (g.tune.cached ? "tuned (cache file)" : "tuned") : "untuned",
                          describe_tune(g.tune.sched).c_str(), g.tune.us_default, g.tune.us_best);
        s += line;
    }
    if (buf && len > 0) {
        const size_t n = std::min<size_t>(s.size(), (size_t)len - 1);
        std::memcpy(buf, s.data(), n);
        buf[n] = 0;
    }
    return (int)s.size() + 1;
}

int minfft_batch(minfft_batch_plan* p, const void* in, size_t in_bytes, void* out, size_t out_bytes) {
    clear_error();
    if (!p) {
        set_error("invalid batch plan");
        return -1;
    }
    // the entries must lie inside the buffers: refused before any copy or launch
    if (in_bytes < p->in_extent || out_bytes < p->out_extent) {
        set_error("minfft_batch: buffer smaller than the entries address (in needs " +
                  std::to_string(p->in_extent) + " bytes, out " + std::to_string(p->out_extent) + ")");
        return -1;
    }
    if (!in || !out) {
        set_error("minfft_batch: null buffer");
        return -1;
    }
    const bool din = is_device_pointer(in), dout = is_device_pointer(out);
    if (din && dout) {
        if (minfft_batch_run(p, in, out, nullptr)) return -1;
        if (hipStreamSynchronize(p->stream) != hipSuccess) {
            set_error("batch execution failed");
            return -1;
        }
        return 0;
    }
    // device staging buffers held by the plan, grown on demand
    auto grow = [](void*& buf, size_t& have, size_t need) -> bool {
        need = std::max<size_t>(need, 1);
        if (have >= need) return true;
        if (buf) (void)hipFree(buf);
        buf = nullptr;
        have = 0;
        if (dev_alloc((void**)&buf, need) != hipSuccess) return false;
        have = need;
        return true;
    };
    if (!grow(p->st_in, p->st_in_bytes, in_bytes) || !grow(p->st_out, p->st_out_bytes, out_bytes)) {
        set_error("hipMalloc staging failed");
        return -1;
    }
    void* di = p->st_in;
    void* dout_p = p->st_out;
    int rc = 0;
    if (hipMemcpyAsync(di, in, in_bytes, hipMemcpyDefault, p->stream) != hipSuccess) rc = -1;
    // regions of `out` not covered by any entry keep their contents
    if (!rc && hipMemcpyAsync(dout_p, out, out_bytes, hipMemcpyDefault, p->stream) != hipSuccess) rc = -1;
    if (!rc) rc = enqueue_batch(p, di, dout_p, p->stream, 1);
    if (!rc && hipMemcpyAsync(out, dout_p, out_bytes, hipMemcpyDefault, p->stream) != hipSuccess) rc = -1;
    if (hipStreamSynchronize(p->stream) != hipSuccess) rc = -1;
    if (rc && !*last_error()) set_error("batch execution failed");
    return rc;
}

}  // extern "C"
